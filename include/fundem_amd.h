/*
 * fundem_amd C-ABI: stateless gfx950 launchers for the level-set DEM timestep.
 *
 * This is the boundary the C++ host API (fundem_amd/include/fundem/LSSolver.h
 * and its containers) calls, in the place of FunDEM's extern "C" launch*
 * layer (kernel/CUDAKernelFunction/ *.cuh, myUtility/mySpatialGrid.h,
 * myUtility/buildHashStartEnd.h). Each declaration names the reference symbol
 * it replaces.
 *
 * Conventions
 *   - Device state is planar SoA: a vector field of n bodies is 3n values
 *     [x0..x(n-1) | y0.. | z0..]; a quaternion field is 4n values
 *     [q0.. | q1.. | q2.. | q3..] (scalar first, as myQua.h:10-15).
 *     Positions are FP64; velocity, angular velocity, orientation, force
 *     and torque are FP32.
 *   - Per-template and per-material constants and grid parameters are FP64.
 *   - All arithmetic that decides an integer outcome (cell ids, contact
 *     membership, bond failure) is FP64 on the device, evaluated in the
 *     reference's operation order without contraction.
 *   - Every launcher is asynchronous on `stream`, allocates nothing and never
 *     synchronises the host, so a step can be captured in a hipGraph. Counts
 *     that the reference read back to the host (numPair) stay on the device.
 *   - Return value: FDM_OK, or a negative FDM_E* code for invalid arguments
 *     or a failed launch. Capacity overflow inside a step is reported through
 *     a device-resident status word (FDM_STATUS_*), read by the host when it
 *     chooses to synchronise.
 */
#ifndef FUNDEM_AMD_H
#define FUNDEM_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDM_OK 0
#define FDM_EINVAL (-1)
#define FDM_ELAUNCH (-2)
#define FDM_ECAPACITY (-3)

/* bits of the device status word */
#define FDM_STATUS_PAIR_OVERFLOW 1u     /* candidate pairs exceeded capacity */
#define FDM_STATUS_RECORD_OVERFLOW 2u   /* contact records exceeded capacity */
#define FDM_STATUS_BUCKET_OVERFLOW 4u   /* per-particle mirror list exceeded capacity */
#define FDM_STATUS_GHOST_OVERFLOW 8u    /* ghost pair list exceeded capacity */
#define FDM_STATUS_STAGING_OVERFLOW 16u /* staged contact geometry of one step (all sets
                                           together) exceeded its workspace */

#define FDM_CHUNK 64         /* surface nodes per culling chunk (one wavefront) */
/* History records are reserved from FDM_REC_POOLS pools, each with its own
 * counter on its own cache line, so concurrent reservations do not all
 * serialise on one address. */
#define FDM_REC_POOLS 64
#define FDM_REC_STRIDE 32
#define FDM_REC_COUNTERS (FDM_REC_POOLS * FDM_REC_STRIDE)
#define FDM_MAX_CHUNKS 64    /* chunks per template: at most 4096 nodes */

typedef struct { double x, y, z; } fdm_double3;
typedef struct { int x, y, z; } fdm_int3;

/* Storage precision of the state a step carries over: velocities, angular
 * velocities, orientations, forces, torques, tangential springs and bond
 * constants, and the templates' surface-node coordinates (positions are
 * always FP64, arithmetic always FP64, SDF values always FP32). The
 * library is built twice from the same sources: libfundem_amd.so stores
 * them in FP32 (the production path, bit for bit the oracle with FP32
 * storage), libfundem_amd_fp64.so (built with FDM_STATE_FP64) in FP64, so
 * that its steps are the reference's all-FP64 arithmetic (bit for bit the
 * oracle with FP64 storage and canonical sums). One history record is the
 * sliding spring of a node-contact and the node's id. */
#ifdef FDM_STATE_FP64
typedef double fdm_real;
typedef struct __attribute__((aligned(16))) {
  double x, y, z;
  int node;
  int pad_;
} fdm_record;
#else
typedef float fdm_real;
typedef struct __attribute__((aligned(16))) {
  float x, y, z;
  int node;
} fdm_record;
#endif

/* One particle or wall template (geometryInfo, LSParticle.h:53-174), FP64. */
typedef struct {
  double radius;         /* max |node| after the centroid shift */
  double invMass;
  double inertia[6];     /* xx yy zz xy xz yz */
  double invInertia[6];
  double origin[3];      /* SDF grid origin in the template frame */
  double invSpacing;
  double lipschitz;      /* bound on |grad| of the trilinear SDF (culling only) */
  double boxReach;       /* max distance from the template frame origin to an SDF box corner */
  long long sdfStart;    /* offset into the concatenated SDF */
  int size[3];           /* SDF grid nodes per axis */
  int nodeStart;         /* offset into the node table (xyz float triples) */
  int nodeCount;
  int chunkStart;        /* offset into the chunk table (float4: centre xyz, radius) */
  int chunkCount;
  int affine;            /* 1: the stored SDF is affine in the box (a plane wall): the
                            trilinear phi at a template-frame point l is at least
                            plane[3] + plane[0..2] . l - planeSlack (culling only) */
  double plane[4];
  double planeSlack;
} fdm_template;

/* One material (materialInfo, LSParticle.h:12-51) after the clamps of
 * LSParticle::addMaterial; dissipation = -ln e / sqrtf(ln^2 e + pi^2) as
 * contactKernel.cuh:50-51 evaluates it. */
typedef struct {
  double kn, ks, mu, e;
  double dissipation;
  double pad_[3];
} fdm_material;

/* Device view of a set of bodies (particles or walls), planar SoA: positions
 * FP64 (a step's displacement, v*dt, can be below one FP32 ulp of the
 * coordinate), everything else fdm_real. */
typedef struct {
  int n;
  double* x;
  fdm_real* v;
  fdm_real* w;
  fdm_real* q;
  fdm_real* F;
  fdm_real* T;
  const int* mat;
  const int* geo;
} fdm_bodies;

/* Geometry tables of one template set. */
typedef struct {
  int num;
  const fdm_template* info;
  const float* sdf;
  const fdm_real* nodes; /* xyz triples after the centroid shift */
  const float* chunks;   /* float4 per chunk */
} fdm_geometry;

/* Uniform grid parameters (spatialGrid, myUtility/mySpatialGrid.h:162-256). */
typedef struct {
  fdm_double3 min, max, invCell;
  fdm_int3 size;
  long long num;
} fdm_grid;

/* Sorted hash table of one grid: hash value per body (by body index), the
 * stably sorted (hash, index) lists and half-open cell ranges. */
typedef struct {
  int* hashValue;      /* [n] by body index */
  unsigned* sortKey;   /* [n] hash + 1, sorted in place */
  int* sortIndex;      /* [n] body index, sorted in place */
  int* cellStart;      /* [grid.num], -1 if empty */
  int* cellEnd;        /* [grid.num] */
  int* prevKey;        /* [n] sorted keys of the previous build, to clear cells */
  int* prevCount;      /* [1] entries in prevKey (-1: clear the whole table) */
  int* validAtBuild;   /* [n] body state at the last list build (0 outside the
                          grid, 1 inside, 2 index rounded up to the grid size);
                          NULL when candidate lists are rebuilt every step */
} fdm_hash_table;


/* Candidate-list reuse. Lists are built with an extra reach `skin` and kept
 * while every body has moved (or, through rotation, moved a point within
 * `rho` of its centre) by less than skin/2 since the build, no body changed
 * grid state, no periodic image appeared or disappeared and no particle
 * wrapped. Each step re-evaluates the exact contact predicate on the kept
 * candidates, so results do not depend on `skin`. The decision is taken on
 * the device every step; a captured step needs no host logic. Control words
 * (int, device): */
#define FDM_CTL_REBUILD 0        /* lists rebuilt this step */
#define FDM_CTL_REBUILD_GHOST 1  /* ghost lists rebuilt this step */
#define FDM_CTL_PENDING 2        /* a particle wrapped: rebuild at the next step */
#define FDM_CTL_VALID_CHANGED 3  /* a body changed grid state this step */
#define FDM_CTL_GHOST_CHANGED 4  /* a periodic image appeared or disappeared */
#define FDM_CTL_WRAPPED 5        /* a particle wrapped this step */
#define FDM_CTL_MAX_RATE 6       /* largest |v| + |w| rho this step (float bits) */
#define FDM_CTL_FORCE 7          /* host-requested rebuild */
#define FDM_CTL_COUNT 8          /* list builds so far */
#define FDM_CTL_GHOST_COUNT 9    /* ghost list builds so far */
#define FDM_CTL_SEQ 10           /* decisions taken so far */
#define FDM_CTL_KEEP 11          /* 1 when this step keeps every list (no rebuild at all) */
#define FDM_CTL_WORDS 16
#define FDM_DECISION_RING 16     /* words of the host-visible decision ring */
typedef struct {
  double skin;         /* extra reach of the candidate lists (0: rebuild every step) */
  double rhoMax;       /* largest lever arm of any body (FP32 orientation rounding) */
  double dt;
  int* ctl;            /* [FDM_CTL_WORDS] */
  double* disp;        /* [1] displacement bound accumulated since the last build */
  int* hostDecision;   /* optional host-visible ring of FDM_DECISION_RING words: decision
                          number k writes word k % FDM_DECISION_RING with
                          (k << 2) | rebuild << 1 | ghost rebuild, last (for a host that
                          launches only the work a step needs, and may queue up to
                          FDM_DECISION_RING - 1 steps ahead of the decisions it has read);
                          NULL to skip */
} fdm_skin;

/* Candidate pairs and tangential history of one interaction set. */
typedef struct {
  int capacity;        /* pair capacity */
  int* count;          /* [n+1] per-A candidate count, scanned in place to CSR starts */
  int* total;          /* [1] number of pairs */
  int* pairA;          /* [capacity] */
  int* pairT;          /* [capacity] stable target id (particle, wall or real id of a ghost) */
  int* pairG;          /* [capacity] ghost index (ghost sets only, else unused) */
  int* recOff;         /* [capacity] first history record of the pair */
  int* recCnt;         /* [capacity] contacts of the pair */
  double* sumF;        /* [3*capacity] per-pair force sum (pair order, node order) */
  double* sumTA;       /* [3*capacity] torque sum about A */
  double* sumTB;       /* [3*capacity] torque sum about the target */
  int recCapacity;     /* records, split into FDM_REC_POOLS equal pools */
  int* recTotal;       /* [FDM_REC_COUNTERS] records used per pool, pool j at
                          recTotal[j * FDM_REC_STRIDE] */
  fdm_record* rec;     /* [recCapacity] (spring xyz, node id), per pair by node */
} fdm_pairs;

/* Per-particle mirror lists: pairs in which the particle is the target. */
typedef struct {
  int capacity;        /* entries per particle */
  int* count;          /* [n] cleared by the reduction after use */
  int* list;           /* [n*capacity] pair indices */
} fdm_mirror;

/* Visible ghosts of one periodic set (Ghost, PeriodicBoundary.h:7-176). */
typedef struct {
  int capacity;
  int* count;          /* [1] */
  int* flag;           /* [n] scan workspace */
  int* realId;         /* [capacity] real particle of each ghost, ascending */
  double* pos;         /* [3*capacity] ghost positions, AoS */
  double* q;           /* [4*capacity] sector images: ghost orientations (NULL for planar sets,
                          whose ghosts keep the real particle's) */
  double* vel;         /* [3*capacity] sector images: ghost velocities (NULL for planar sets) */
  fdm_hash_table table; /* hashValue: the reference's ghost cell id (stencil rule);
                           sortKey: 1 + id of the clamped geometric cell (list builds) */
  int* cellBox;        /* [6] min/max occupied sort cell per axis */
  int* atBuild;        /* [n] per particle: visibility bits (X 1, Y 2, XY 4) at the
                          last ghost list build; NULL when rebuilt every step */
} fdm_ghosts;

/* Bonds (VBondedInteraction, VBondedInteraction.h:61-566), fdm_real constants. */
typedef struct {
  int n;
  const int* master;
  const int* slave;
  int* active;
  int* breakStep;      /* step index at which the bond failed, -1 while intact */
  const fdm_real* coef;   /* [n*9]: B1 B2 B3 B4 R L0 tensile cohesion friction */
  const fdm_real* local;  /* [n*24]: n1i n2i n3i pi n1j n2j n3j pj (xyz each) */
  double* outF;        /* [3n] bond force on master */
  double* outTi;       /* [3n] torque on master */
  double* outTj;       /* [3n] torque on slave */
  int* applied;        /* [n] 1 if the bond acted this step */
  double* maxNormalStress; /* [n] */
  double* maxShearStress;  /* [n] */
  const int* partStart;    /* [np+1] CSR of bonds per particle (ascending bond id) */
  const int* partList;     /* bond id * 2 + role (0 master, 1 slave) */
  const int* stepIndex;    /* [1] device step counter, recorded in breakStep */
  const int* gate;         /* run only when *gate != 0 (NULL: always) */
} fdm_bonds;

/* Periodic configuration: planar x / y (PeriodicBoundaryXY2D,
 * PeriodicBoundary.h:178-317), or the quarter-plane sector about the z axis
 * through the grid's min corner (PeriodicBoundarySector, :319-426; then x and
 * y are 0). Sector ghost sets 0, 1, 2 hold the R90, R180 and R270 images. The
 * sector's rotations are given by fdm_sectorPeriodic, per image k (0: +pi/2,
 * 1: pi, 2: -pi/2): the angle about z, rotateVector's coefficients
 * A = sin(t)/t and B = (1 - cos t)/t^2 (t = |angle|; myVec.h:143-181) and
 * rotateQuaternion's cos and sin of t/2 (myQua.h:131-175), evaluated on the
 * host as the reference's host code does, so that the device and a host
 * restatement rotate with the same constants. */
typedef struct {
  int x, y;
  int sector;
  int pad_;
  double angle[3], A[3], B[3], halfCos[3], halfSin[3];
} fdm_periodic;
/* Fill p for the sector boundary (sector = 1, x = y = 0, rotation constants). */
void fdm_sectorPeriodic(fdm_periodic* p);

/* Independent systems advanced together (no counterpart in the reference,
 * which runs one system per solver). Every body carries an instance id; each
 * instance keeps its own coordinates, particle grid, wall grid, periodic
 * flags and materials, so it reproduces its own plain run bit for bit. The
 * instances' cells are laid end to end in one global cell space: a body's
 * sort key is 1 + (first global cell of its instance) + (its cell id in its
 * instance's grid), so one sort and one table serve every instance. count = 0
 * means a single system described by the launchers' own grid arguments. */
typedef struct {
  int count;
  const int* particleInstance;  /* [particles] */
  const int* wallInstance;      /* [walls]; walls are stored by instance */
  const int* wallStart;         /* [count+1] walls of instance k: wallStart[k] .. wallStart[k+1] */
  const fdm_grid* pgrid;        /* [count] particle grid of each instance */
  const fdm_grid* wgrid;        /* [count] wall grid of each instance */
  const long long* pcellOffset; /* [count] first global particle cell of each instance */
  const long long* wcellOffset; /* [count] first global wall cell of each instance */
  const fdm_periodic* periodic; /* [count] */
} fdm_instances;

/* ---------------------------------------------------------------------- */
/* Integration                                                            */

/* Replaces launchParticleVelocityAngularVelocityIntegration
 * (particleIntegrationKernel.cuh:4): half-kick of v and w with step h. */
int fdm_launchParticleVelocityAngularVelocityIntegration(fdm_bodies bodies, fdm_geometry geo,
                                                         fdm_double3 gravity, double h,
                                                         hipStream_t stream);

/* The tutorials' external-force hooks (apps/tutorial/globalDamping.cu:4-63),
 * for LSSolver::addExternalForceTorque overrides on this layout. Global
 * damping: F -= c |F| v/|v| and T -= c |T| w/|w| per particle (a zero
 * velocity gives a zero direction). Constant force: adds F to the force of
 * particle `index`, and then T to its force as well, as the tutorial's kernel
 * does (its torque is left alone). */
int fdm_launchAddGlobalDampingForceTorque(fdm_bodies particles, double dampingCoefficient,
                                          hipStream_t stream);
int fdm_launchAddGlobalConstantForceTorque(fdm_bodies particles, fdm_double3 constantForce,
                                           fdm_double3 constantTorque, int index,
                                           hipStream_t stream);
/* Global damping with the tutorials' argument list (launchAddGlobalDampingForceTorque,
 * apps/tutorial/globalDamping.cuh:6-15) on this layout: force, torque, velocity and
 * angular velocity of `num` bodies as fdm_real planar SoA (LSParticle::force(), torque(),
 * velocity(), angularVelocity()), on gridD workgroups of blockD threads
 * (LSParticle::gridDim(), blockDim()) striding over the bodies. Same results as
 * fdm_launchAddGlobalDampingForceTorque. */
int fdm_launchAddGlobalDampingForceTorqueArrays(fdm_real* force, fdm_real* torque,
                                               const fdm_real* velocity,
                                               const fdm_real* angularVelocity,
                                               double dampingCoefficient, size_t num,
                                               size_t gridD, size_t blockD, hipStream_t stream);

/* Replaces launchParticlePositionOrientationIntegration
 * (particleIntegrationKernel.cuh:23): drift of x and q over dt. */
int fdm_launchParticlePositionOrientationIntegration(fdm_bodies bodies, double dt,
                                                     hipStream_t stream);

/* Fused first half of a step (LSSolver::integration1stHalf, LSSolver.h:1033-1070,
 * plus calculateHashKernel, mySpatialGrid.cu:4-26): optional half-kick with
 * the stored F,T, drift, then the body's cell id and sort key/value. With
 * `skin` (may be NULL) it also folds the body's motion rate into
 * ctl[FDM_CTL_MAX_RATE] and flags ctl[FDM_CTL_VALID_CHANGED] when the body's
 * grid state differs from table.validAtBuild. */
int fdm_launchKickDriftHash(fdm_bodies bodies, fdm_geometry geo, fdm_double3 gravity, double dt,
                            int kick, int drift, fdm_grid grid, fdm_hash_table table,
                            const fdm_skin* skin, hipStream_t stream);
/* No counterpart in the reference (which rebuilds every list every step).
 * Decides, on the device, whether this step rebuilds the candidate lists:
 * accumulates the displacement bound and sets ctl[FDM_CTL_REBUILD] and
 * ctl[FDM_CTL_REBUILD_GHOST] (after fdm_launchPeriodicPrepare, when
 * periodic); a wrap seen by the prepare stage also rebuilds the next step. */
int fdm_launchSkinDecide(const fdm_skin* skin, hipStream_t stream);
/* No counterpart in the reference: the first phase of a step in two
 * launches. fdm_launchStepBegin: kick (when `kick`), drift (when `drift`)
 * and cell id (table.hashValue) of every particle, drift and cell id of
 * every wall, the periodic visibility of every particle against its state at
 * the last ghost list build (the flags are stored only where ghostFlag[k] is
 * not NULL: list builds compute their own), leaving per-workgroup partials in
 * `workspace` (fdm_stepBeginWorkspaceBytes(particles + walls) bytes). The
 * sort lists are left to fdm_launchCommitBodyState, which only steps that
 * rebuild lists run.
 * fdm_launchStepDecide (same arguments, after it): clears the given counter
 * arrays and takes the list decision of fdm_launchSkinDecide from the
 * partials. */
typedef struct {
  fdm_bodies particles;
  fdm_geometry pgeo;
  fdm_grid pgrid;
  fdm_hash_table ptable;
  fdm_bodies walls;            /* n may be 0 */
  fdm_geometry wgeo;
  fdm_grid wgrid;
  fdm_hash_table wtable;
  fdm_double3 gravity;
  double dt;
  int kick, drift;
  fdm_periodic periodic;
  int* ghostFlag[3];           /* per particle visibility flag of sets X, Y, XY */
  int* ghostAtBuild;           /* [n] visibility bits at the last ghost build */
  fdm_instances inst;          /* count 0: one system (pgrid, wgrid, periodic above) */
  int* clearCounters[8];       /* arrays of FDM_REC_COUNTERS ints set to 0 (NULL: skipped) */
  const int* gate;             /* both launches run only when *gate != 0 (NULL: always) */
} fdm_step_begin_args;
size_t fdm_stepBeginWorkspaceBytes(size_t bodies);
int fdm_launchStepBegin(const fdm_step_begin_args* args, const fdm_skin* skin, void* workspace,
                        size_t workspaceBytes, hipStream_t stream);
int fdm_launchStepDecide(const fdm_step_begin_args* args, const fdm_skin* skin, void* workspace,
                         size_t workspaceBytes, hipStream_t stream);
/* No counterpart in the reference. When ctl[FDM_CTL_REBUILD] is set, records
 * every body's grid state in table.validAtBuild and writes the table's
 * unsorted lists for the build (sortKey = cell id + 1, sortIndex = body
 * index; with instances, the key in the global cell space). */
int fdm_launchCommitBodyState(fdm_bodies bodies, fdm_grid grid, fdm_hash_table table,
                              const fdm_skin* skin, hipStream_t stream);
/* The same for bodies of several instances (isWall selects the wall grids). */
int fdm_launchCommitBodyStateInstances(fdm_bodies bodies, const fdm_instances* inst, int isWall,
                                       fdm_hash_table table, const fdm_skin* skin,
                                       hipStream_t stream);
/* No counterpart in the reference. When *gate is 0, copies the candidate
 * lists (count, total, pairA/T/G) of `from` into `to`, so a step without a
 * rebuild sees the lists of the last build. */
int fdm_launchCarryPairs(const fdm_pairs* from, const fdm_pairs* to, int n, const int* gate,
                         hipStream_t stream);

/* ---------------------------------------------------------------------- */
/* Spatial hash                                                           */

/* Bytes of workspace fdm_radixSortPairs needs for n keys. */
size_t fdm_radixSortWorkspaceBytes(size_t n);

/* Stable LSD radix sort of (key, value) by key, keys < 2^keyBits, in place.
 * Replaces thrust::sort_by_key as called by buildHashStartEnd
 * (buildHashStartEnd.cu:108-111). */
int fdm_radixSortPairs(unsigned* keys, int* values, size_t n, int keyBits, void* workspace,
                       size_t workspaceBytes, hipStream_t stream);

/* Replaces buildHashStartEnd (myUtility/buildHashStartEnd.h:32): sorts the
 * table's keys (hash + 1) with the body indices and writes half-open cell
 * ranges; cells of the previous build are reset first. */
int fdm_buildHashStartEnd(fdm_hash_table table, int n, fdm_grid grid, void* workspace,
                          size_t workspaceBytes, hipStream_t stream);
/* The same, as a no-op unless *gate is nonzero (gate may be NULL: always). */
int fdm_buildHashStartEndGated(fdm_hash_table table, int n, fdm_grid grid, const int* gate,
                               void* workspace, size_t workspaceBytes, hipStream_t stream);

/* No counterpart in the reference. fdm_buildHashStartEnd without a workspace
 * argument, for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the sort's workspace is a fixed device array inside the library, one per
 * GPU, held until the process ends. The call allocates nothing. The caller
 * still passes the table (inputs and outputs). Returns FDM_ECAPACITY for
 * n > FDM_NOALLOC_MAX_BODIES. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race;
 * calls ordered on one stream do not. Results equal fdm_buildHashStartEnd's
 * bit for bit. */
#define FDM_NOALLOC_MAX_BODIES (1 << 21)
int fdm_buildHashStartEndNoAlloc(fdm_hash_table table, int n, fdm_grid grid, hipStream_t stream);

/* Replaces launchUpdateSpatialGridHashStartEnd (myUtility/mySpatialGrid.h:33):
 * cell id of every body (calculateHashKernel) then buildHashStartEnd. */
int fdm_launchUpdateSpatialGridHashStartEnd(fdm_hash_table table, const double* position, int n,
                                            fdm_grid grid, void* workspace, size_t workspaceBytes,
                                            hipStream_t stream);

/* Replaces buildPrefixSum (myUtility/buildHashStartEnd.h:62): inclusive scan
 * of n ints (in may equal out). */
size_t fdm_scanWorkspaceBytes(size_t n);
int fdm_buildPrefixSum(int* prefixSum, const int* count, size_t n, void* workspace,
                       size_t workspaceBytes, hipStream_t stream);

/* ---------------------------------------------------------------------- */
/* Periodic boundaries                                                    */

/* Replace launchUpdatePositionOutOfBoundaryXD / YD (mySpatialGrid.h:51/59). */
int fdm_launchUpdatePositionOutOfBoundaryXD(double* position, int n, fdm_grid grid,
                                            hipStream_t stream);
int fdm_launchUpdatePositionOutOfBoundaryYD(double* position, int n, fdm_grid grid,
                                            hipStream_t stream);
/* Replace launchCalculateGhostPositionXD / YD / XYD (mySpatialGrid.h:76/90/104).
 * axisMask: 1 = X, 2 = Y, 3 = XY. Ghost positions are FP64 AoS. */
int fdm_launchCalculateGhostPosition(double* ghostPosition, const double* position, int n,
                                     fdm_grid grid, int axisMask, hipStream_t stream);
/* Replaces launchUpdateGhostSpatialGridHashStartEnd (mySpatialGrid.h:144). */
int fdm_launchUpdateGhostSpatialGridHashStartEnd(fdm_hash_table table, const double* ghostPosition,
                                                 int n, fdm_grid grid, void* workspace,
                                                 size_t workspaceBytes, hipStream_t stream);

/* Fused periodic stage of one step (PeriodicBoundaryXY2D::updateGhostSpatialGrid,
 * PeriodicBoundary.h:221-295): wrap, then the visible ghosts of every active
 * set, compacted in particle order and hashed into per-set tables. */
int fdm_launchPeriodicGhosts(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                             fdm_ghosts* sets /* [3]: X, Y, XY */, unsigned* status,
                             void* workspace, size_t workspaceBytes, hipStream_t stream);
/* The same stage split for candidate-list reuse (no counterpart in the
 * reference). Prepare: visibility flags, flagging ctl[FDM_CTL_WRAPPED] and
 * ctl[FDM_CTL_GHOST_CHANGED] against sets[].atBuild; it runs before
 * fdm_launchSkinDecide. Build: the ghost sets from those flags (as
 * fdm_launchPeriodicGhosts), recording the visibility bits; it runs when
 * *gate != 0 (gate NULL: always). Refresh: positions and cell ids of the
 * kept ghosts; it runs when *gate == 0 (gate NULL: always). */
int fdm_launchPeriodicPrepare(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                              fdm_ghosts* sets, const fdm_skin* skin, hipStream_t stream);
int fdm_launchPeriodicBuild(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                            fdm_ghosts* sets, const int* gate, unsigned* status, void* workspace,
                            size_t workspaceBytes, hipStream_t stream);
/* fdm_launchPeriodicBuild for several instances: `grid` only gives the
 * global cell count (grid.num); `periodic` says which sets exist. */
int fdm_launchPeriodicBuildInstances(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                     const fdm_instances* inst, fdm_ghosts* sets, unsigned* status,
                                     void* workspace, size_t workspaceBytes, hipStream_t stream);
/* The wrap of the stored positions (launchUpdatePositionOutOfBoundaryXD/YD,
 * mySpatialGrid.cu:28-56, in updateGhostSpatialGrid's order). Build and
 * BuildInstances compute the ghosts from the wrapped positions but leave the
 * stored positions unwrapped: the reference detects particle and wall contacts
 * with the positions before the wrap and ghost contacts and forces after it
 * (LSSolver.h:1168-1178), so the step calls this after its detection. Runs
 * when *gate != 0 (gate NULL: always). */
int fdm_launchPeriodicCommitWrap(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                                 const fdm_instances* inst, const int* gate, hipStream_t stream);
int fdm_launchPeriodicRefresh(fdm_bodies particles, fdm_grid grid, fdm_periodic periodic,
                              fdm_ghosts* sets, const int* gate, hipStream_t stream);

/* ---------------------------------------------------------------------- */
/* Contact detection and contact forces                                   */

#define FDM_SET_PARTICLE 0
#define FDM_SET_WALL 1
#define FDM_SET_GHOST 2

/* Broad phase of one interaction set: candidate pairs (A, target) sorted by
 * (A, target), every pair that can hold a node contact under the reference's
 * detection rules (LSParticleContactDetectionKernel.cu:35-123) included.
 * Replaces the counting half of launchBuildLevelSetBoundaryNodeInteractions1st
 * / FixedParticleInteractions1st (LSParticleContactDetectionKernel.cuh:79/148)
 * together with buildPrefixSum / updateNumPair (SolidInteraction.h:92-120). */
typedef struct {
  int set;                 /* FDM_SET_* */
  fdm_bodies particles;
  fdm_geometry pgeo;
  /* targets: particles (set 0), walls (set 1) or ghosts (set 2) */
  fdm_bodies walls;
  fdm_geometry wgeo;
  const fdm_ghosts* ghosts;
  fdm_grid grid;           /* grid of the targets */
  fdm_hash_table table;    /* hash table of the targets */
  fdm_pairs pairs;
  fdm_instances inst;      /* count 0: one system on `grid` */
  double skin;             /* extra reach (candidate-list reuse), 0 for exact lists */
  double maxReach;         /* largest target reach: particle radius (set 0) or
                              SDF box reach (set 2) over all templates */
  const int* gate;         /* run only when *gate != 0 (NULL: always) */
  fdm_periodic periodic;   /* ghost sets: the wrap applied to A's stored position (ghost
                              contacts are detected after the wrap, fdm_launchPeriodicCommitWrap) */
} fdm_broadphase_args;
int fdm_launchBroadPhase(const fdm_broadphase_args* args, void* workspace, size_t workspaceBytes,
                         hipStream_t stream);

/* Narrow phase fused with the contact law, for every candidate pair:
 * node sweep in wave64 against the target's SDF, the reference's contact
 * predicate, history remap keyed by (A, target, node) from the previous
 * step's pairs, linear spring-dashpot with Coulomb-capped sliding spring,
 * and per-pair force/torque sums in node order. Replaces
 * launchBuildLevelSetBoundaryNode*Interactions2nd
 * (LSParticleContactDetectionKernel.cuh:108/180) and
 * launchAdd{,Fixed,Ghost}LevelSetParticleContactForceTorque
 * (contactKernel.cuh:201/234/275). */
typedef struct {
  int set;
  fdm_bodies particles;
  fdm_geometry pgeo;
  const fdm_material* pmat;
  fdm_bodies walls;
  fdm_geometry wgeo;
  const fdm_material* wmat;
  const fdm_ghosts* ghosts;
  fdm_grid grid;               /* grid of the targets (for the stencil rule) */
  const int* targetHash;       /* hash value per target */
  fdm_grid pgrid;              /* particle grid (periodic wrap of force arms) */
  fdm_periodic periodic;       /* wrap applied to positions used in forces */
  fdm_pairs pairs;             /* this step */
  fdm_pairs prev;              /* previous step (history); prev.count is its CSR by A */
  fdm_mirror mirror;           /* append pairs with contacts to the target's list */
  int ghostMask;               /* ghost sets: 1 X, 2 Y, 3 XY */
  int ghostFromReal;           /* ghost sets: take each ghost's position and cell id from
                                  its real particle (valid when no particle wrapped this
                                  step) instead of ghosts->pos / ghosts->table.hashValue */
  const int* gate;             /* fdm_launchNarrowPhaseSets: run only when *gate != 0
                                  (NULL: always; taken from sets[0]) */
  fdm_instances inst;          /* count 0: one system on grid / pgrid / periodic */
  double dt;
  int updateSprings;           /* 0: keep remapped springs (history filtering) */
  unsigned* status;
  /* optional per-record diagnostics (NULL to skip): contact point, normal, overlap */
  double* dbgPoint;
  double* dbgNormal;
  double* dbgOverlap;
} fdm_narrowphase_args;
int fdm_launchNarrowPhase(const fdm_narrowphase_args* args, hipStream_t stream);
/* Scratch of fdm_launchNarrowPhaseSets. */
typedef struct {
  int* workList;               /* [8 * workCapacity] or NULL: per candidate pair, the header
                                  the sweep writes for the contact law (record count, first
                                  record, first staged contact, A, target, previous record
                                  range), so that the law starts from one load */
  int workCapacity;            /* at least the sum of the sets' pair capacities */
  int* workCount;              /* unused (may be NULL); kept for the layout */
  int* workHit;                /* [workCapacity] first staged contact of each candidate pair */
  double* hitGeom;             /* [8 * hitCapacity] staged contacts */
  int hitCapacity;             /* staged contacts, split into FDM_REC_POOLS pools */
  int* hitCount;               /* [FDM_REC_COUNTERS] staged contacts per pool */
  const struct fdm_tuning_s* tuning; /* launch shapes (NULL: the process default) */
} fdm_narrow_workspace;
/* No counterpart in the reference: the narrow phase of up to five sets at
 * once, in two kernels over the candidate pairs of all sets (set k's pairs
 * follow those of sets 0..k-1). The first (one wave per candidate pair)
 * tests whether the pair's bodies can still touch at this step's positions
 * (the broad-phase tests without the skin), gives a rejected pair empty
 * record ranges, and otherwise sweeps the surface nodes, stages every
 * contact's point, normal and overlap in hitGeom and reserves the pair's
 * records; the second (one thread per candidate pair) applies the contact
 * law to the staged contacts in node order and writes the records and the
 * pair sums. Shared fields (bodies, geometry, materials, pgrid, periodic,
 * dt, updateSprings, status, dbg*) come from sets[0]. Unlike
 * fdm_launchNarrowPhase it clears no counter: the sets' record counters
 * (pairs.recTotal) and the hitCount pools must be zero at launch. Candidate
 * g reserves its records and staged contacts from pool g % FDM_REC_POOLS. */
int fdm_launchNarrowPhaseSets(const fdm_narrowphase_args* sets, int nsets,
                              const fdm_narrow_workspace* ws, hipStream_t stream);

/* Replaces launchAddBondedForceTorque (contactKernel.cuh:311): VBond forces,
 * stresses and one-way failure; results per bond. */
int fdm_launchBondForces(fdm_bodies particles, fdm_bonds bonds, hipStream_t stream);

/* Second half of the broad phase: writes the candidate pairs counted and
 * scanned by fdm_launchBroadPhase, sorted by target within each A. */
int fdm_launchBroadPhaseWrite(const fdm_broadphase_args* args, unsigned* status,
                              hipStream_t stream);

/* Add one to a device counter (the step index a captured graph replays). */
int fdm_incrementCounter(int* counter, hipStream_t stream);

/* Deterministic reduction (replaces the double atomicAdds of
 * contactKernel.cu:120-123,203-204,284-290,367-370 and the memsets of
 * LSSolver.h:919-923), fused with the second half-kick
 * (LSSolver::integration2ndHalf, LSSolver.h:1072-1089). Per particle, records
 * are added in this order: particle pairs by partner (as target, then as A),
 * wall pairs by wall, each ghost set X, Y, XY by partner real id (as A before
 * as target), bonds by id. kick = 0 only writes F and T. */
typedef struct {
  fdm_bodies particles;
  fdm_geometry pgeo;
  const fdm_pairs* pp;
  const fdm_mirror* ppMirror;
  const fdm_pairs* pw;
  const fdm_pairs* ghost[3];   /* NULL when the set is inactive */
  const fdm_mirror* ghostMirror[3];
  const fdm_bonds* bonds;      /* NULL when there are none */
  fdm_double3 gravity;
  double h;
  int kick;
  int* stepCounter;    /* incremented by one after the reduction (NULL: skip) */
  const int* gate;     /* run only when *gate != 0 (NULL: always) */
  const struct fdm_tuning_s* tuning; /* launch shapes (NULL: the process default) */
} fdm_reduce_args;
int fdm_launchReduceForceTorque(const fdm_reduce_args* args, hipStream_t stream);

/* Fill n ints with a value (a kernel, not a memset node, so it replays
 * correctly inside captured graphs). */
int fdm_fillInt(int* ptr, size_t n, int value, hipStream_t stream);

/* No counterpart in the reference: launch shapes, picked by the autotuner
 * (fundem::LSAutotuner) for one solver. Results do not depend on them. A
 * launcher uses the shapes its arguments point to (fdm_narrow_workspace.tuning,
 * fdm_reduce_args.tuning), or the process default when that pointer is NULL.
 * fdm_setTuning changes the process default (a zero field keeps the current
 * value); fdm_checkTuning fills the zero fields of a caller's shapes from the
 * default and validates them. Graphs captured before a change keep the shapes
 * they were captured with. */
typedef struct fdm_tuning_s {
  int sweepWaves;       /* node-sweep waves (kept pairs in flight) per workgroup: 1, 2 or 4 */
  int sweepWavesPerCU;  /* node-sweep grid size, in waves per CU: 4..128 (default 32) */
  int lawBlock;         /* contact-law threads per workgroup: 64, 128 or 256 */
  int reduceBlock;      /* reduction threads per workgroup: a multiple of 64 up to 1024 */
  int lawMode;          /* contact law of the narrow phase: 1 in a kernel of its own (one
                           thread per candidate pair, contacts staged in the workspace),
                           2 in the sweep's wave (one contact per lane) */
  int nodesPerLane;     /* how a sweep wave walks a pair's surviving 64-node chunks: 1 one
                           chunk per iteration (one node per lane), 2 two chunks per
                           iteration (128 nodes in flight), 3 the nodes that pass the tests
                           before the SDF fetch collected first, then up to 64 of their SDF
                           cells fetched at once; lawMode 1 only */
  int wallGroups;       /* wall pairs four per wave in a kernel of their own (nodesPerLane 1,
                           lawMode 1): 0 automatic (wall sets sized for at least 65536
                           pairs), 1 always, 2 never */
} fdm_tuning;
int fdm_setTuning(const fdm_tuning* t);
void fdm_getTuning(fdm_tuning* t);
int fdm_checkTuning(fdm_tuning* t);

/* Device properties used by the launchers (number of CUs); 0 on success. */
int fdm_deviceInfo(int* numCU, int* warpSize);

/* ------------------------------------------------------------------ adjoint
 * No counterpart in the reference: the reverse (adjoint) timestep behind
 * fundem::LSAdjoint, for gradients of a scalar outcome of the final state
 * with respect to the initial state and the contact parameters. Supported:
 * particle-particle, particle-wall (walls with prescribed motion) and
 * periodic ghost contacts, several systems in one solver; no bonds. */

/* Contacts of consecutive steps, block by block: block b holds a step's
 * contacts set by set (particle pairs, wall pairs, ghost X, Y, XY pairs),
 * each set in canonical order (A, target, node), in the slot
 * [b * maxBlockContacts, (b + 1) * maxBlockContacts) of every array. */
typedef struct {
  int capacity;        /* contacts over all slots: at least maxBlocks * maxBlockContacts */
  int maxBlocks;
  int* blockStart;     /* [maxBlocks] first contact of the block (b * maxBlockContacts) */
  int* blockPart;      /* [6 * maxBlocks] block-local start of each set, then the count */
  int* blockCount;     /* [maxBlocks] all its contacts */
  int* A;              /* [capacity] */
  int* T;              /* [capacity] target particle or wall */
  int* node;           /* [capacity] node of A */
  fdm_real* spring;    /* [3 * capacity] tangential spring after the step, xyz per contact */
  int* prev;           /* [capacity] the same (A, target, node) in the previous block, or -1 */
  unsigned* status;    /* FDM_STATUS_RECORD_OVERFLOW when a block does not fit */
} fdm_contact_tape;

/* Workspace of the tape routines for blocks of at most maxBlockContacts
 * contacts and sets whose pair capacities sum to pairCapacity. */
size_t fdm_tapeWorkspaceBytes(int maxBlockContacts, long long pairCapacity);
/* Write the contacts of the latest step as block `block`. sets: the five
 * sets' current lists and results in tape order (particle, wall, ghost X,
 * Y, XY), NULL for an absent set (sets[0] is required). prevBlock: the block
 * the step's history came from, -1 for none. */
int fdm_launchTapeAppend(const fdm_pairs* const* sets, fdm_contact_tape tape, int block, int prevBlock,
                         int maxBlockContacts, void* ws, size_t wsBytes, hipStream_t stream);
/* Write block `block` into the sets' lists and results (the generations a
 * following step reads as its history; their count arrays hold
 * nParticles + 1 entries). The following step must rebuild its lists. */
int fdm_launchTapeRestore(fdm_contact_tape tape, int block, const fdm_pairs* const* sets,
                          int nParticles, int maxBlockContacts, void* ws, size_t wsBytes,
                          hipStream_t stream);

/* One step in reverse. The state before the step (x, q, v, w and the F, T
 * of the previous step) and after it (x, q, F, T) come from the tape; the
 * adjoint arrays hold the adjoint of the state after the step on entry and
 * of the state before it on return. */
typedef struct {
  fdm_bodies before;           /* particles at the start of the step */
  fdm_bodies after;            /* particles at its end */
  fdm_geometry pgeo;
  const fdm_material* pmat;
  fdm_bodies walls;            /* walls during the step (after their drift) */
  fdm_geometry wgeo;
  const fdm_material* wmat;
  fdm_double3 gravity;
  double dt;
  fdm_grid pgrid;              /* particle grid (ghost periods) */
  fdm_instances inst;          /* count 0: one system on pgrid */
  fdm_contact_tape tape;
  int block;                   /* the step's contacts */
  int prevBlock;               /* the block its history came from (-1: none) */
  int maxBlockContacts;
  double* xBar;                /* [3n] planar */
  double* qBar;                /* [4n] */
  double* vBar;                /* [3n] */
  double* wBar;                /* [3n] */
  double* FBar;                /* [3n] adjoint of the force the step starts with */
  double* TBar;                /* [3n] */
  double* springBar;           /* [3 * tape.capacity] adjoint of each contact's spring: read for
                                  `block`, written for prevBlock */
  double* paramBar;            /* [4n] accumulated: each particle's share of the kn, ks, mu,
                                  zeta adjoints, for its own material */
  int* wallMat;                /* [2n] the wall materials each particle's wall contacts used
                                  (at most two; -1 before the first) */
  double* wallShare;           /* [4n] accumulated mu, zeta adjoints of those two materials */
  unsigned* status;            /* FDM_STATUS_BUCKET_OVERFLOW: more than 64 particle-pair contacts
                                  on one target in a step, or a third wall material */
} fdm_adjoint_step_args;
size_t fdm_adjointWorkspaceBytes(int nParticles, int maxBlockContacts);
/* Copy the state arrays (x, q, v, w, F, T) of n bodies from src to dst. */
int fdm_launchCopyBodies(fdm_bodies dst, fdm_bodies src, hipStream_t stream);
int fdm_launchAdjointStep(const fdm_adjoint_step_args* args, void* ws, size_t wsBytes,
                          hipStream_t stream);

/* ---------------------------------------------------------------------- */
/* Mesh signed distance (input preparation)                               */

/* Replaces the grid loop of LevelSetObject::Base::buildGrid
 * (LevelSetObjectGenerator.h:186-223) for a TriangleMeshParticle
 * (LevelSetObjectGenerator.h:888-914): the level-set value of a closed
 * triangle mesh at every node origin + (i, j, k) * spacing of a
 * size[0] x size[1] x size[2] grid (x fastest), written to sdf. triangles:
 * [9 * numTriangles] device FP64 vertex coordinates a b c per triangle.
 * Distance: the minimum over triangles (closestPointOnTriangle, :454-495);
 * sign: the solid-angle sum over triangles (:758-800); positiveInside flips
 * it (setLevelSetPositiveInside). origin and size are host arrays. */
int fdm_launchMeshSignedDistanceGrid(const double* triangles, int numTriangles,
                                     const double* origin, double spacing, const int* size,
                                     int positiveInside, double* sdf, hipStream_t stream);
/* TriangleMeshParticle::evaluateImplicitFunctionValueAndGradient
 * (LevelSetObjectGenerator.h:888-914) at numPoints device points [3*numPoints]:
 * value [numPoints] and gradient [3*numPoints]. */
int fdm_launchMeshSignedDistancePoints(const double* triangles, int numTriangles,
                                       const double* points, int numPoints, int positiveInside,
                                       double* value, double* grad, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
