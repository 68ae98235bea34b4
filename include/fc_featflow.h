/* FEATFLOW's Fortran entry points for the particle state of rigid bodies
 * moved by the particle step (inshape3dcore/fortrancppinterface/
 * cppwrapper_linux.h, the names gfortran gives the calls of the CFD code).
 * They act on the coupling bound with fc_featflow_bind (fc_dem.h); every
 * argument is passed by reference, as Fortran does.  They return nothing,
 * like the reference's: on a bad id or with no coupling bound they leave
 * their outputs unchanged and set fc_last_error(). */
#ifndef FC_FEATFLOW_H
#define FC_FEATFLOW_H

#ifdef __cplusplus
extern "C" {
#endif

/* update_particle_state_ (cppwrapper_linux.h:28-42): position, velocity,
 * Euler angles and angular velocity of body *iid (the angles and angular
 * velocities are 0: the particle step moves spheres without rotation). */
void update_particle_state_(double* px, double* py, double* pz, double* vx, double* vy,
                            double* vz, double* ax, double* ay, double* az, double* avx,
                            double* avy, double* avz, int* iid);
/* getpos_ / getvel_ (cppwrapper_linux.h:717-725). */
void getpos_(double* dx, double* dy, double* dz, int* iID);
void getvel_(double* dx, double* dy, double* dz, int* iID);
/* setpositionid_ / setvelocityid_ (cppwrapper_linux.h:702-716). */
void setpositionid_(double* dx, double* dy, double* dz, int* iID);
void setvelocityid_(double* dx, double* dy, double* dz, int* iID);
/* settimestep_ (cppwrapper_linux.h:368-372). */
void settimestep_(double* dTime);
/* getnumparticles_ (cppwrapper_linux.h:413-417): the number of bodies of
 * the coupling (the reference counts its boundary body too). */
void getnumparticles_(int* nParts);
/* getradius_ (cppwrapper_linux.h:418-422): the spheres' radius. */
void getradius_(double* drad, int* iid);
/* startcollisionpipeline_ (cppwrapper_linux.h:453-457): one time step of
 * the bound coupling in 10 particle updates. */
void startcollisionpipeline_(void);

#ifdef __cplusplus
}
#endif

#endif /* FC_FEATFLOW_H */
