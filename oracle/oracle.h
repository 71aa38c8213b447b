/*
 * TEST INFRASTRUCTURE ONLY - never linked into or called by the product library.
 *
 * Serial FP64 CPU restatement of FunDEM's level-set DEM timestep, used as the
 * parity oracle for the gfx950 HIP path. Arrays are AoS doubles: a vector is
 * (x,y,z), a quaternion (q0,q1,q2,q3) scalar first, a symmetric matrix
 * (xx,yy,zz,xy,xz,yz), exactly the reference's double3 / quaternion / symMatrix
 * layouts (kernel/CUDAKernelFunction/myUtility/myVec.h, myQua.h:10-15,
 * myMat.h:15-27).
 *
 * Two switches select what is restated:
 *   store_fp32     0: every stored value stays FP64 (the reference's storage).
 *                  1: values are rounded to FP32 exactly where the device path
 *                     stores FP32 (velocity, angular velocity and orientation
 *                     after each integration stage, tangential springs, summed
 *                     force/torque, uploaded SDF/nodes/bond constants);
 *                     positions stay FP64 as on the device.
 *   canonical_fold 0: contact and bond forces are added one contact at a time in
 *                     contact-list order, a serial execution of the reference's
 *                     double atomicAdds (contactKernel.cu:120-123,203-204,284-290,
 *                     367-370).
 *                  1: forces are summed per interaction pair in node order, then
 *                     per particle in the canonical order documented in
 *                     oracle.cpp (fold_canonical), the order the device's
 *                     deterministic reduction uses.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

/* ---- stage restatements (stateless) ---------------------------------- */

/* spatialGrid::set (myUtility/mySpatialGrid.h:212-240). max_io may be
 * adjusted in place exactly as the reference does. */
void oracle_grid_set(double* min3, double* max3, double cell_size, int* size3, double* inv_cell3);

/* calculateHashKernel (myUtility/mySpatialGrid.cu:4-26). pos: n*3 doubles. */
void oracle_hash(int n, const double* pos, const double* min3, const double* max3,
                 const double* inv_cell3, const int* size3, int* hash_out);

/* calculateGhostHashKernel (myUtility/mySpatialGrid.cu:214-240), including the
 * z-against-max.y comparison. */
void oracle_ghost_hash(int n, const double* pos, const double* min3, const double* max3,
                       const double* inv_cell3, const int* size3, int* hash_out);

/* buildHashStartEnd (myUtility/buildHashStartEnd.cu:69-137): stable sort of
 * (hash, index) by hash, then half-open start/end per cell (-1 if empty). */
void oracle_sort_start_end(int n, int* hash_inout, int* index_out, long long num_grid,
                           int* start_out, int* end_out);

/* thrust::inclusive_scan as used by buildPrefixSum (buildHashStartEnd.cu:139-178). */
void oracle_inclusive_scan(long long n, const int* count, int* prefix_out);

/* ---- whole-solver restatement ---------------------------------------- */

typedef struct OracleSim OracleSim;

OracleSim* osim_create(int store_fp32, int canonical_fold);
void osim_destroy(OracleSim* s);

/* LSParticle::addMaterial (LSParticle.h:229-240) with its clamps. */
int osim_add_material(OracleSim* s, double kn, double ks, double mu, double e);
/* LSSolver::addWallMaterial (LSSolver.h:175-183): kn = ks = 0. */
int osim_add_wall_material(OracleSim* s, double mu, double e);
/* Additive damping-ratio materials: the dashpots use zeta directly; e =
 * exp(-pi zeta / sqrt(1 - zeta^2)) only decides which material of a pair
 * applies (the one with the smaller e, as for restitution materials). */
int osim_add_material_damping(OracleSim* s, double kn, double ks, double mu, double zeta);
int osim_add_wall_material_damping(OracleSim* s, double mu, double zeta);
/* LSParticle::addGeometry (LSParticle.h:242-290). sdf: nx*ny*nz (x fastest),
 * nodes: nnodes*3. Returns the geometry id or -1. */
int osim_add_geometry(OracleSim* s, double mass, const double* inertia6, const double* origin3,
                      double spacing, const int* size3, const double* sdf, const double* nodes,
                      int nnodes);
/* LSSolver::addParticleGeometry(density, ...) (LSSolver.h:470-574): mass,
 * centroid and inertia from the smooth Heaviside of the SDF; nodes and origin
 * shifted by -centroid. local_offset_out receives the centroid. */
int osim_add_geometry_density(OracleSim* s, double density, const double* origin3, double spacing,
                              const int* size3, const double* sdf, const double* nodes, int nnodes,
                              double* local_offset_out3);
/* LSSolver::addWallGeometry (LSSolver.h:194-210). */
int osim_add_wall_geometry(OracleSim* s, const double* origin3, double spacing, const int* size3,
                           const double* sdf, const double* nodes, int nnodes);
/* LSParticle::addParticle (LSParticle.h:292-342): orientation normalised. */
int osim_add_particle(OracleSim* s, const double* x3, const double* v3, const double* w3,
                      const double* q4, int material, int geometry);
int osim_add_wall(OracleSim* s, const double* x3, const double* v3, const double* w3,
                  const double* q4, int material, int geometry);
/* VBondedInteraction::add (VBondedInteraction.h:74-195) via
 * LSSolver::addBondedInteraction (LSSolver.h:306-337). */
int osim_add_bond(OracleSim* s, int master, int slave, const double* point3, const double* normal3,
                  double radius, double length, double youngs, double poisson, double tensile,
                  double cohesion, double friction);
void osim_set_periodic(OracleSim* s, int x_on, int y_on);
/* PeriodicBoundarySector (PeriodicBoundary.h:319-426): the quarter-plane
 * sector x >= min.x, y >= min.y about the z axis through min; ghost sets
 * 2, 3, 4 hold the R90, R180 and R270 images. Exclusive with the planar sets. */
void osim_set_sector(OracleSim* s, int on);
/* The tutorials' external forces (apps/tutorial/globalDamping.cu): global
 * damping with coefficient c (c < 0: off) and the constant-force kernel on one
 * particle (index < 0: off), applied between the forces and the second
 * half-kick. */
void osim_set_global_damping(OracleSim* s, double c);
void osim_set_constant_force(OracleSim* s, int index, const double* F3, const double* T3);
/* Additive: overrides the particle grid's 2*max radius cell size (<= 0 restores it). */
void osim_set_cell_size(OracleSim* s, double cell_size);
/* LSSolver::upload (LSSolver.h:692-703) then the pre-loop grid/contact build of
 * the compute* variants (LSSolver.h:1097-1100,1146-1149,1172-1175,1201-1204). */
int osim_initialize(OracleSim* s, const double* min3, const double* max3, const double* gravity3,
                    double dt);
/* One iteration of the compute* loop body (LSSolver.h:1101-1110 and variants). */
int osim_step(OracleSim* s, int nsteps);
/* Brute-force particle-particle detection (no grid) for the current state:
 * returns the number of contacts and fills (A, B, k) in canonical order. */
long long osim_bruteforce_pp(OracleSim* s, int* a_out, int* b_out, int* k_out, long long cap);

int osim_num_particles(const OracleSim* s);
int osim_num_walls(const OracleSim* s);
int osim_num_bonds(const OracleSim* s);
void osim_get_state(const OracleSim* s, double* x, double* v, double* w, double* q, double* f,
                    double* t);
void osim_set_state(OracleSim* s, const double* x, const double* v, const double* w,
                    const double* q, const double* f, const double* t);
void osim_get_walls(const OracleSim* s, double* x, double* q);
void osim_get_geometry(const OracleSim* s, int geo, double* radius, double* inv_mass,
                       double* inertia6, double* inv_inertia6, double* origin3, double* inv_h);
void osim_get_grid(const OracleSim* s, int walls, double* min3, double* max3, double* inv_cell3,
                   int* size3);
/* set: 0 particle-particle, 1 particle-wall, 2 ghost X, 3 ghost Y, 4 ghost XY.
 * Contacts in canonical order (A, target, k). Any output pointer may be NULL.
 * Returns the contact count (outputs truncated at cap). */
long long osim_get_contacts(const OracleSim* s, int set, int* a, int* target, int* k,
                            double* spring3, double* point3, double* normal3, double* overlap,
                            long long cap);
/* Hash arrays of the last grid build: hash (sorted), index, start/end. */
void osim_get_hash(const OracleSim* s, int walls, int* sorted_hash, int* index, int* start,
                   int* end);
void osim_get_bonds(const OracleSim* s, int* active, double* max_normal_stress,
                    double* max_shear_stress);
/* Set one entry of the tangential history of a contact set, looked up by key;
 * returns 1 if the key was present. Used by tests to seed histories. */
int osim_set_spring(OracleSim* s, int set, int a, int target, int k, const double* spring3);
/* Ghost set `set` (0 X, 1 Y, 2 XY; sector: R90, R180, R270) of the latest
 * step, one entry per particle: the ghost position (the particle's own
 * position where it has no image) and its ghost cell id
 * (calculateGhostHashKernel; -1 where the ghost is invisible). Returns the
 * entry count (0 before the first periodic step). */
int osim_get_ghosts(const OracleSim* s, int set, double* pos, int* hash);

/* Mesh level sets (meshsdf.cpp; LevelSetObjectGenerator.h:122-224,454-914).
 * v: [3*nv] vertices, f: [3*nf] triangles. Points: value and gradient at np
 * points (value 1, gradient 0 for an invalid mesh). Grid: Base::buildGrid's
 * placement (origin3, size3) and, when sdf is not null, its values; returns
 * the node count (0 for an invalid mesh or spacing). */
void oracle_mesh_points(const double* v, int nv, const int* f, int nf, const double* pts, int np,
                        int positiveInside, double* value, double* grad);
long long oracle_mesh_grid(const double* v, int nv, const int* f, int nf, double h, int padding,
                           int positiveInside, double* origin3, int* size3, double* sdf);

/* Level-set sampling of the detection kernels (LSParticleContactDetectionKernel.cu:72-119,
 * .cuh:4-77) on one grid of size3 nodes (x fastest) at origin3 with the given spacing:
 * trilinear value and the Q1 gradient per cell at np local-frame points; inside[i] = 0
 * (value and gradient 0) where the point's cell is not inside the grid. */
void oracle_sample_level_set(const double* sdf, const int* size3, const double* origin3,
                             double spacing, const double* pts, int np, double* value,
                             double* grad, int* inside);

#ifdef __cplusplus
}
#endif
