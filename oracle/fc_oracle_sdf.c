/*
 * fc_oracle_sdf.c - CPU restatement of the signed-distance surrogate MLP
 * (sdf_mlp.hpp / sdf_mlp.hip of fullc0ntact_amd): forward and backward in
 * binary64, and a forward that rounds weights and activations to bfloat16
 * as the MFMA inference does.  No counterpart in the reference.
 *
 * TEST INFRASTRUCTURE ONLY (see fc_oracle.h).  Parameter layout per body:
 * W1[32][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
 * (W[input][output]); the loss is 1/n sum (y - label/halfExtent)^2.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "fc_oracle.h"

#define NF 16
#define NIN 32
#define NH 64
#define OW1 0
#define OB1 (OW1 + NIN * NH)
#define OW2 (OB1 + NH)
#define OB2 (OW2 + NH * NH)
#define OW3 (OB2 + NH)
#define OB3 (OW3 + NH * NH)
#define OW4 (OB3 + NH)
#define OB4 (OW4 + NH)
#define NP (OB4 + 1)

int orc_sdf_num_params(void) { return NP; }

static const double TWO_PI = 6.283185307179586476925286766559;

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double silu_grad(double z) {
  const double s = 1.0 / (1.0 + exp(-z));
  return s * (1.0 + z * (1.0 - s));
}

/* Round to the nearest bfloat16, ties to even (as float). */
static float bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}

static void features(const float* freq, const float* frame, const float* q, double* f) {
  const double x = ((double)q[0] - frame[0]) / frame[3];
  const double y = ((double)q[1] - frame[1]) / frame[3];
  const double z = ((double)q[2] - frame[2]) / frame[3];
  for (int k = 0; k < NF; ++k) {
    const double t = freq[3 * k] * x + freq[3 * k + 1] * y + freq[3 * k + 2] * z;
    f[k] = sin(TWO_PI * t);
    f[NF + k] = cos(TWO_PI * t);
  }
}

/* Forward of one query; fills the pre-activations when z1..z3 are given. */
static double forward(const float* P, const double* f, double* z1, double* z2, double* z3) {
  double a1[NH], a2[NH];
  for (int o = 0; o < NH; ++o) {
    double s = P[OB1 + o];
    for (int i = 0; i < NIN; ++i) s += f[i] * P[OW1 + i * NH + o];
    if (z1) z1[o] = s;
    a1[o] = silu(s);
  }
  for (int o = 0; o < NH; ++o) {
    double s = P[OB2 + o];
    for (int i = 0; i < NH; ++i) s += a1[i] * P[OW2 + i * NH + o];
    if (z2) z2[o] = s;
    a2[o] = silu(s);
  }
  double y = P[OB4];
  for (int o = 0; o < NH; ++o) {
    double s = P[OB3 + o];
    for (int i = 0; i < NH; ++i) s += a2[i] * P[OW3 + i * NH + o];
    if (z3) z3[o] = s;
    y += silu(s) * P[OW4 + o];
  }
  return y;
}

void orc_sdf_forward(const float* params, const float* freq, const float* frame,
                     const float* pts, uint32_t n, double* out) {
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < (long)n; ++i) {
    double f[NIN];
    features(freq, frame, pts + 3 * i, f);
    out[i] = forward(params, f, NULL, NULL, NULL) * frame[3];
  }
}

/* Gradient of L = 1/n sum (y - label/h)^2 (binary64, sequential sums). */
void orc_sdf_loss_grad(const float* params, const float* freq, const float* frame,
                       const float* pts, const float* labels, uint32_t n, double* grad,
                       double* loss) {
  const float* P = params;
  memset(grad, 0, sizeof(double) * NP);
  double L = 0.0;
  for (uint32_t s = 0; s < n; ++s) {
    double f[NIN], z1[NH], z2[NH], z3[NH], a1[NH], a2[NH];
    features(freq, frame, pts + 3 * s, f);
    const double y = forward(P, f, z1, z2, z3);
    for (int o = 0; o < NH; ++o) {
      a1[o] = silu(z1[o]);
      a2[o] = silu(z2[o]);
    }
    const double err = y - (double)labels[s] / frame[3];
    L += err * err;
    const double dy = 2.0 * err / n;
    double d3[NH], d2[NH], d1[NH];
    grad[OB4] += dy;
    for (int o = 0; o < NH; ++o) {
      grad[OW4 + o] += silu(z3[o]) * dy;
      d3[o] = dy * P[OW4 + o] * silu_grad(z3[o]);
    }
    for (int o = 0; o < NH; ++o) {
      grad[OB3 + o] += d3[o];
      for (int i = 0; i < NH; ++i) grad[OW3 + i * NH + o] += a2[i] * d3[o];
    }
    for (int i = 0; i < NH; ++i) {
      double s2 = 0.0;
      for (int o = 0; o < NH; ++o) s2 += d3[o] * P[OW3 + i * NH + o];
      d2[i] = s2 * silu_grad(z2[i]);
    }
    for (int o = 0; o < NH; ++o) {
      grad[OB2 + o] += d2[o];
      for (int i = 0; i < NH; ++i) grad[OW2 + i * NH + o] += a1[i] * d2[o];
    }
    for (int i = 0; i < NH; ++i) {
      double s1 = 0.0;
      for (int o = 0; o < NH; ++o) s1 += d2[o] * P[OW2 + i * NH + o];
      d1[i] = s1 * silu_grad(z1[i]);
    }
    for (int o = 0; o < NH; ++o) {
      grad[OB1 + o] += d1[o];
      for (int i = 0; i < NIN; ++i) grad[OW1 + i * NH + o] += f[i] * d1[o];
    }
  }
  *loss = L / n;
}

/* The BF16 inference arithmetic: weights (W1..W3) and every matmul input
 * (features, hidden activations) rounded to bfloat16, products accumulated
 * in binary64; biases, w4 and the last layer in full precision. */
void orc_sdf_forward_bf16(const float* params, const float* freq, const float* frame,
                          const float* pts, uint32_t n, double* out) {
  long q;
#pragma omp parallel for schedule(static)
  for (q = 0; q < (long)n; ++q) {
    const float* P = params;
    float x[3];
    for (int a = 0; a < 3; ++a) x[a] = (pts[3 * q + a] - frame[a]) / frame[3];
    float f[NIN];
    for (int k = 0; k < NF; ++k) {
      const float t = freq[3 * k] * x[0] + freq[3 * k + 1] * x[1] + freq[3 * k + 2] * x[2];
      f[k] = bf16((float)sin(TWO_PI * (double)t));
      f[NF + k] = bf16((float)cos(TWO_PI * (double)t));
    }
    float h[NH], g[NH];
    for (int o = 0; o < NH; ++o) {
      double s = 0.0;
      for (int i = 0; i < NIN; ++i) s += (double)f[i] * bf16(P[OW1 + i * NH + o]);
      h[o] = bf16((float)silu((double)(float)s + P[OB1 + o]));
    }
    for (int o = 0; o < NH; ++o) {
      double s = 0.0;
      for (int i = 0; i < NH; ++i) s += (double)h[i] * bf16(P[OW2 + i * NH + o]);
      g[o] = bf16((float)silu((double)(float)s + P[OB2 + o]));
    }
    double y = P[OB4];
    for (int o = 0; o < NH; ++o) {
      double s = 0.0;
      for (int i = 0; i < NH; ++i) s += (double)g[i] * bf16(P[OW3 + i * NH + o]);
      y += silu((double)(float)s + P[OB3 + o]) * P[OW4 + o];
    }
    out[q] = y * frame[3];
  }
}
