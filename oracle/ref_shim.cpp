// C entry point over FullC0ntact's own point-triangle distance
// (inshape3dcore/distance/distancetriangle.cpp), compiled from the reference
// sources by oracle/ref.mk into oracle/_ref/libref_distance.so.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's Eberly
// distance against the reference's implementation.
#include <distancetriangle.h>
#include <triangle3.h>

extern "C" double ref_point_triangle_distance(const double* q, const double* a, const double* b,
                                              const double* c, double* closest) {
  using i3d::Vector3;
  i3d::Triangle3<double> tri(Vector3<double>(a[0], a[1], a[2]), Vector3<double>(b[0], b[1], b[2]),
                             Vector3<double>(c[0], c[1], c[2]));
  i3d::CDistancePointTriangle<double> d(tri, Vector3<double>(q[0], q[1], q[2]));
  const double r = d.ComputeDistance();
  closest[0] = d.m_vClosestPoint1.x;
  closest[1] = d.m_vClosestPoint1.y;
  closest[2] = d.m_vClosestPoint1.z;
  return r;
}
