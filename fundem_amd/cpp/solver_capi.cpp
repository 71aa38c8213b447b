// C binding of the C++ host API (LSSolver), for callers without C++: the
// Python package, tests and benchmark bind it with ctypes. Every entry returns
// 0 or a negative code; fdmx_last_error() gives the message.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../include/fundem/LSAdjoint.h"
#include "../include/fundem/LSAutotune.h"
#include "../include/fundem/LSSolver.h"
#include "../include/fundem/LevelSetObjectGenerator.h"
#include "../include/fundem/OBJLoader.h"

namespace {

thread_local std::string g_error;

// Exposes the protected density path of addParticleGeometry.
class BoundSolver : public fundem::LSSolver {
 public:
  explicit BoundSolver(int device) : fundem::LSSolver("fundem", 0, 256, device) {}
  // The tutorials' external forces (apps/tutorial/LSParticleFalling.cpp:26-45):
  // global damping with coefficient c (c < 0: off) and a constant force on
  // one particle (index < 0: off).
  void setGlobalDamping(double c) {
    damping_ = c;
    resetHookDetection();
  }
  void setConstantForce(int index, double3 F, double3 T) {
    constIndex_ = index;
    constF_ = F;
    constT_ = T;
    resetHookDetection();
  }
  void addExternalForceTorque(const double time) override {
    if (damping_ < 0.0 && constIndex_ < 0) {
      fundem::LSSolver::addExternalForceTorque(time);
      return;
    }
    fdm_bodies b = getLSParticle().bodies();
    if (damping_ >= 0.0)
      fundem::fdmCheck(fdm_launchAddGlobalDampingForceTorque(b, damping_, stream()), "damping");
    if (constIndex_ >= 0)
      fundem::fdmCheck(fdm_launchAddGlobalConstantForceTorque(b, {constF_.x, constF_.y, constF_.z},
                                                              {constT_.x, constT_.y, constT_.z},
                                                              constIndex_, stream()),
                       "constant force");
  }

  void addGeometryDensity(double3& off, double density, double3 origin, double h, int3 size,
                          const std::vector<double>& sdf, const std::vector<double3>& nodes,
                          const std::vector<int3>& tris) {
    addParticleGeometry(off, density, origin, h, size, sdf, nodes, tris);
  }
 private:
  double damping_ = -1.0;
  int constIndex_ = -1;
  double3 constF_{0.0, 0.0, 0.0}, constT_{0.0, 0.0, 0.0};
};

// tris: ntri triangles of node indices (the surface mesh the frame files
// draw); may be null with ntri = 0.
static std::vector<int3> i3v(const int* t, int n) {
  std::vector<int3> out((size_t)std::max(n, 0));
  for (int k = 0; k < n; ++k) out[k] = make_int3(t[3 * k], t[3 * k + 1], t[3 * k + 2]);
  return out;
}
template <class F>
int guarded(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  } catch (...) {
    g_error = "unknown error";
    return -1;
  }
}

double3 d3(const double* p) { return make_double3(p[0], p[1], p[2]); }
std::vector<double3> d3v(const double* p, int n) {
  std::vector<double3> v(n);
  for (int i = 0; i < n; ++i) v[i] = d3(p + 3 * i);
  return v;
}

}  // namespace

extern "C" {

const char* fdmx_last_error() { return g_error.c_str(); }

void* fdmx_create(int device) {
  try {
    return new BoundSolver(device);
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void fdmx_destroy(void* s) { delete static_cast<BoundSolver*>(s); }

int fdmx_add_material(void* s, double kn, double ks, double mu, double e) {
  return guarded([&] { static_cast<BoundSolver*>(s)->addParticleMaterial(kn, ks, mu, e); });
}
int fdmx_add_wall_material(void* s, double mu, double e) {
  return guarded([&] { static_cast<BoundSolver*>(s)->addWallMaterial(mu, e); });
}
int fdmx_add_material_damping(void* s, double kn, double ks, double mu, double zeta) {
  return guarded(
      [&] { static_cast<BoundSolver*>(s)->addParticleMaterialDampingRatio(kn, ks, mu, zeta); });
}
int fdmx_add_wall_material_damping(void* s, double mu, double zeta) {
  return guarded([&] { static_cast<BoundSolver*>(s)->addWallMaterialDampingRatio(mu, zeta); });
}
int fdmx_num_materials(void* s, int* particle, int* wall) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    *particle = (int)S->getLSParticle().numMaterial();
    *wall = (int)S->getWall().numMaterial();
  });
}

// Contact geometry of the latest step's contacts of a set (host recomputed):
// returns the count (outputs truncated at cap) or -1.
long long fdmx_contact_geometry(void* s, int set, int* A, int* T, int* k, double* point,
                                double* normal, double* overlap, long long cap) {
  long long n = -1;
  guarded([&] {
    const auto c = static_cast<BoundSolver*>(s)->contactGeometry(set);
    n = (long long)c.size();
    for (long long i = 0; i < n && i < cap; ++i) {
      if (A) A[i] = c[i].A;
      if (T) T[i] = c[i].target;
      if (k) k[i] = c[i].node;
      if (point) { point[3 * i] = c[i].point.x; point[3 * i + 1] = c[i].point.y; point[3 * i + 2] = c[i].point.z; }
      if (normal) { normal[3 * i] = c[i].normal.x; normal[3 * i + 1] = c[i].normal.y; normal[3 * i + 2] = c[i].normal.z; }
      if (overlap) overlap[i] = c[i].overlap;
    }
  });
  return n;
}
int fdmx_create_bonds_from_contacts(void* s, double R, double L, double E, double nu, double st,
                                    double coh, double mub) {
  return guarded([&] {
    static_cast<BoundSolver*>(s)->createBondsFromLSInteractions(R, L, E, nu, st, coh, mub);
  });
}
// The frame files of one output frame (LSSolver::output) under dir.
int fdmx_write_vtu(void* s, const char* dir, long long frame, long long step, double time) {
  return guarded([&] { static_cast<BoundSolver*>(s)->writeVTU(dir, (size_t)frame, (size_t)step, time); });
}
long long fdmx_num_bonds(void* s) {
  return (long long)static_cast<BoundSolver*>(s)->getLSParticleBondedInteraction().numPair();
}

// Autotuner (fundem::LSAutotuner): out9 = skin, cell size, sweep waves,
// sweep waves per CU, law block, reduce block, ms per step, law mode,
// nodes per lane; log
// receives one line per timed candidate (truncated at logCap).
int fdmx_autotune(void* s, int trialSteps, int tuneCell, double* out7, char* log, int logCap) {
  return guarded([&] {
    fundem::LSAutotuner t(*static_cast<BoundSolver*>(s));
    const fundem::TuneChoice c = t.run((size_t)std::max(trialSteps, 4), tuneCell != 0);
    if (out7) {
      out7[0] = c.skin;
      out7[1] = c.cellSize;
      out7[2] = c.launch.sweepWaves;
      out7[3] = c.launch.sweepWavesPerCU;
      out7[4] = c.launch.lawBlock;
      out7[5] = c.launch.reduceBlock;
      out7[6] = c.msPerStep;
      out7[7] = c.launch.lawMode;
      out7[8] = c.launch.nodesPerLane;
    }
    if (log && logCap > 0) {
      std::string all;
      for (const auto& l : c.trials) all += l + "\n";
      std::strncpy(log, all.c_str(), (size_t)logCap - 1);
      log[logCap - 1] = 0;
    }
  });
}

// Bytes per stored state value of this build (4: FP32, 8: FP64 storage).
int fdmx_state_bytes() { return (int)sizeof(fdm_real); }

// Launch shapes of one solver, seven fields in fdm_tuning order (zero
// fields: the process default).
int fdmx_set_launch_tuning(void* s, const int* t5) {
  return guarded([&] {
    fdm_tuning t{t5[0], t5[1], t5[2], t5[3], t5[4], t5[5], t5[6]};
    static_cast<BoundSolver*>(s)->setLaunchTuning(t);
  });
}
int fdmx_get_launch_tuning(void* s, int* t5) {
  return guarded([&] {
    const fdm_tuning& t = static_cast<BoundSolver*>(s)->launchTuning();
    t5[0] = t.sweepWaves;
    t5[1] = t.sweepWavesPerCU;
    t5[2] = t.lawBlock;
    t5[3] = t.reduceBlock;
    t5[4] = t.lawMode;
    t5[5] = t.nodesPerLane;
    t5[6] = t.wallGroups;
  });
}
int fdmx_set_lookahead(void* s, int steps) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setLookahead(steps); });
}

// Adjoint (fundem::LSAdjoint) of a solver; the solver must outlive it.
void* fdmx_adjoint_create(void* s) {
  try {
    return new fundem::LSAdjoint(*static_cast<BoundSolver*>(s));
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void fdmx_adjoint_destroy(void* a) { delete static_cast<fundem::LSAdjoint*>(a); }
int fdmx_adjoint_set_contacts(void* a, int perParticle) {
  return guarded([&] { static_cast<fundem::LSAdjoint*>(a)->setContactsPerParticle(perParticle); });
}
int fdmx_adjoint_forward(void* a, long long steps, long long every) {
  return guarded([&] { static_cast<fundem::LSAdjoint*>(a)->forward((size_t)steps, (size_t)every); });
}
// Adjoints of the final state in (AoS: 3 or 4 per particle); gradients out:
// initial state (AoS) and per material (particle: kn, ks, mu, zeta; wall:
// mu, zeta). Any output may be NULL.
int fdmx_adjoint_backward(void* a, int n, const double* xb, const double* qb, const double* vb,
                          const double* wb, double* x0, double* q0, double* v0, double* w0,
                          double* kn, double* ks, double* mu, double* zeta, double* wmu,
                          double* wzeta) {
  return guarded([&] {
    std::vector<double3> X(n), V(n), Wv(n);
    std::vector<quaternion> Q(n);
    for (int i = 0; i < n; ++i) {
      X[i] = d3(xb + 3 * i);
      V[i] = d3(vb + 3 * i);
      Wv[i] = d3(wb + 3 * i);
      Q[i] = make_quaternion(qb[4 * i], qb[4 * i + 1], qb[4 * i + 2], qb[4 * i + 3]);
    }
    const fundem::AdjointGradients g = static_cast<fundem::LSAdjoint*>(a)->backward(X, Q, V, Wv);
    for (int i = 0; i < n; ++i) {
      if (x0) { x0[3 * i] = g.position[i].x; x0[3 * i + 1] = g.position[i].y; x0[3 * i + 2] = g.position[i].z; }
      if (v0) { v0[3 * i] = g.velocity[i].x; v0[3 * i + 1] = g.velocity[i].y; v0[3 * i + 2] = g.velocity[i].z; }
      if (w0) {
        w0[3 * i] = g.angularVelocity[i].x;
        w0[3 * i + 1] = g.angularVelocity[i].y;
        w0[3 * i + 2] = g.angularVelocity[i].z;
      }
      if (q0) {
        q0[4 * i] = g.orientation[i].q0;
        q0[4 * i + 1] = g.orientation[i].q1;
        q0[4 * i + 2] = g.orientation[i].q2;
        q0[4 * i + 3] = g.orientation[i].q3;
      }
    }
    for (size_t m = 0; m < g.kn.size(); ++m) {
      if (kn) kn[m] = g.kn[m];
      if (ks) ks[m] = g.ks[m];
      if (mu) mu[m] = g.mu[m];
      if (zeta) zeta[m] = g.dampingRatio[m];
    }
    for (size_t m = 0; m < g.wallMu.size(); ++m) {
      if (wmu) wmu[m] = g.wallMu[m];
      if (wzeta) wzeta[m] = g.wallDampingRatio[m];
    }
  });
}
int fdmx_add_geometry(void* s, double mass, const double* I6, const double* o3, double h,
                      const int* sz, const double* sdf, const double* nodes, int nn, const int* tris,
                      int ntri) {
  return guarded([&] {
    const size_t ns = size_t(sz[0]) * sz[1] * sz[2];
    static_cast<BoundSolver*>(s)->addParticleGeometry(
        mass, symMatrix(I6[0], I6[1], I6[2], I6[3], I6[4], I6[5]), d3(o3), h,
        make_int3(sz[0], sz[1], sz[2]), std::vector<double>(sdf, sdf + ns), d3v(nodes, nn),
        i3v(tris, ntri));
  });
}
int fdmx_add_geometry_density(void* s, double density, const double* o3, double h, const int* sz,
                              const double* sdf, const double* nodes, int nn, double* off3,
                              const int* tris, int ntri) {
  return guarded([&] {
    const size_t ns = size_t(sz[0]) * sz[1] * sz[2];
    double3 off;
    static_cast<BoundSolver*>(s)->addGeometryDensity(off, density, d3(o3), h,
                                                     make_int3(sz[0], sz[1], sz[2]),
                                                     std::vector<double>(sdf, sdf + ns), d3v(nodes, nn),
                                                     i3v(tris, ntri));
    off3[0] = off.x;
    off3[1] = off.y;
    off3[2] = off.z;
  });
}
int fdmx_add_wall_geometry(void* s, const double* o3, double h, const int* sz, const double* sdf,
                           const double* nodes, int nn, const int* tris, int ntri) {
  return guarded([&] {
    const size_t ns = size_t(sz[0]) * sz[1] * sz[2];
    static_cast<BoundSolver*>(s)->addWallGeometry(d3(o3), h, make_int3(sz[0], sz[1], sz[2]),
                                                  std::vector<double>(sdf, sdf + ns), d3v(nodes, nn),
                                                  i3v(tris, ntri));
  });
}
int fdmx_add_particles(void* s, int n, const double* x, const double* v, const double* w,
                       const double* q, const int* mat, const int* geo) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    for (int i = 0; i < n; ++i)
      S->addParticle(d3(x + 3 * i), d3(v + 3 * i), d3(w + 3 * i),
                     make_quaternion(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]), mat[i],
                     geo[i]);
  });
}
int fdmx_add_wall(void* s, const double* x, const double* v, const double* w, const double* q,
                  int mat, int geo) {
  return guarded([&] {
    static_cast<BoundSolver*>(s)->addWall(d3(x), d3(v), d3(w),
                                          make_quaternion(q[0], q[1], q[2], q[3]), mat, geo);
  });
}
int fdmx_add_bond(void* s, int i, int j, const double* pt, const double* nrm, double R, double L,
                  double E, double nu, double st, double coh, double mub) {
  return guarded([&] {
    static_cast<BoundSolver*>(s)->addBondedInteraction(i, j, d3(pt), d3(nrm), R, L, E, nu, st, coh, mub);
  });
}
int fdmx_set_periodic(void* s, int x, int y) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    S->removePeriodicBoundary();
    if (x) S->addPeriodicBoundaryXD();
    if (y) S->addPeriodicBoundaryYD();
  });
}
int fdmx_set_global_damping(void* s, double c) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setGlobalDamping(c); });
}
int fdmx_set_constant_force(void* s, int index, const double* F3, const double* T3) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setConstantForce(index, d3(F3), d3(T3)); });
}
int fdmx_set_sector(void* s, int on) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    S->removePeriodicBoundary();
    if (on) S->addPeriodicBoundarySector();
  });
}
int fdmx_add_system(void* s, const double* mn, const double* mx, int px, int py) {
  int k = -1;
  const int r = guarded([&] {
    k = static_cast<BoundSolver*>(s)->addSystem(d3(mn), d3(mx), px != 0, py != 0);
  });
  return r != 0 ? r : k;
}
int fdmx_set_cell_size(void* s, double cs) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setCellSize(cs); });
}
int fdmx_set_capacity(void* s, int pairs, int records) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setContactCapacity(pairs, records); });
}
int fdmx_set_skin(void* s, double skin) {
  return guarded([&] { static_cast<BoundSolver*>(s)->setNeighborSkin(skin); });
}
int fdmx_list_builds(void* s, long long* builds, long long* ghostBuilds, double* skin) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    const auto b = S->listBuilds();
    if (builds) *builds = b.first;
    if (ghostBuilds) *ghostBuilds = b.second;
    if (skin) *skin = S->neighborSkin();
  });
}
int fdmx_pair_totals(void* s, int* out5) {
  return guarded([&] {
    const auto t = static_cast<BoundSolver*>(s)->pairTotals();
    for (int k = 0; k < 5; ++k) out5[k] = t[k];
  });
}
int fdmx_use_graph(void* s, int on) {
  return guarded([&] { static_cast<BoundSolver*>(s)->useGraph(on != 0); });
}
int fdmx_set_timing(void* s, int on) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    S->setStageTiming(on != 0);
    S->resetStageTiming();
  });
}
int fdmx_initialize(void* s, const double* mn, const double* mx, const double* g, double dt) {
  return guarded([&] { static_cast<BoundSolver*>(s)->initialize(d3(mn), d3(mx), d3(g), dt); });
}
int fdmx_advance(void* s, long long n) {
  return guarded([&] { static_cast<BoundSolver*>(s)->advance((size_t)n); });
}
int fdmx_sync(void* s) {
  return guarded([&] {
    fundem::hipCheck(hipStreamSynchronize(static_cast<BoundSolver*>(s)->stream()), "sync");
  });
}
int fdmx_status(void* s, unsigned* out) {
  return guarded([&] { *out = static_cast<BoundSolver*>(s)->status(); });
}
int fdmx_get_state(void* s, double* x, double* v, double* w, double* q, double* f, double* t) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    S->downloadState();
    auto& P = S->getLSParticle();
    const size_t n = P.num();
    for (size_t i = 0; i < n; ++i) {
      const double3 a = P.positionHostRef()[i], b = P.velocityHostRef()[i],
                    c = P.angularVelocityHostRef()[i], d = P.forceHostRef()[i],
                    e = P.torqueHostRef()[i];
      const quaternion r = P.orientationHostRef()[i];
      if (x) { x[3 * i] = a.x; x[3 * i + 1] = a.y; x[3 * i + 2] = a.z; }
      if (v) { v[3 * i] = b.x; v[3 * i + 1] = b.y; v[3 * i + 2] = b.z; }
      if (w) { w[3 * i] = c.x; w[3 * i + 1] = c.y; w[3 * i + 2] = c.z; }
      if (f) { f[3 * i] = d.x; f[3 * i + 1] = d.y; f[3 * i + 2] = d.z; }
      if (t) { t[3 * i] = e.x; t[3 * i + 1] = e.y; t[3 * i + 2] = e.z; }
      if (q) { q[4 * i] = r.q0; q[4 * i + 1] = r.q1; q[4 * i + 2] = r.q2; q[4 * i + 3] = r.q3; }
    }
  });
}
long long fdmx_num_particles(void* s) {
  return (long long)static_cast<BoundSolver*>(s)->getLSParticle().num();
}
// Contacts of the latest step in canonical order; returns the count (outputs
// truncated at cap) or -1.
long long fdmx_get_contacts(void* s, int set, int* A, int* T, int* k, double* spring, long long cap) {
  long long n = -1;
  guarded([&] {
    auto c = static_cast<BoundSolver*>(s)->contacts(set);
    n = (long long)c.size();
    for (long long i = 0; i < n && i < cap; ++i) {
      if (A) A[i] = c[i].A;
      if (T) T[i] = c[i].target;
      if (k) k[i] = c[i].node;
      if (spring) {
        spring[3 * i] = c[i].spring.x;
        spring[3 * i + 1] = c[i].spring.y;
        spring[3 * i + 2] = c[i].spring.z;
      }
    }
  });
  return n;
}
int fdmx_get_bonds(void* s, int* active, int* breakStep, double* sigma, double* tau) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    auto& B = S->getLSParticleBondedInteraction();
    B.finalize(S->stream());
    const size_t n = B.numPair();
    if (n == 0) return;
    auto sg = B.maxNormalStress(S->stream());
    auto ta = B.maxShearStress(S->stream());
    for (size_t i = 0; i < n; ++i) {
      if (active) active[i] = B.activatedHostRef()[i];
      if (breakStep) breakStep[i] = B.breakStepHostRef()[i];
      if (sigma) sigma[i] = sg[i];
      if (tau) tau[i] = ta[i];
    }
  });
}
// Hash table of the last grid build (walls != 0: wall grid): keys (hash+1)
// and indices in sorted order, and the cell ranges.
int fdmx_get_hash(void* s, int walls, int* hashValue, unsigned* sortedKey, int* sortedIndex,
                  int* start, int* end) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    auto& P = walls ? S->getWall() : S->getLSParticle();
    auto& t = P.table();
    const size_t n = P.num();
    const size_t ng = (size_t)P.grid().num;
    if (hashValue) t.hashValue.download(hashValue, n, S->stream());
    if (sortedKey) t.sortKey.download(sortedKey, n, S->stream());
    if (sortedIndex) t.sortIndex.download(sortedIndex, n, S->stream());
    if (start) t.cellStart.download(start, ng, S->stream());
    if (end) t.cellEnd.download(end, ng, S->stream());
    fundem::hipCheck(hipStreamSynchronize(S->stream()), "hash download");
  });
}
int fdmx_get_grid(void* s, int walls, double* mn, double* mx, double* inv, int* size) {
  return guarded([&] {
    auto* S = static_cast<BoundSolver*>(s);
    const fdm_grid& g = walls ? S->getWall().grid() : S->getLSParticle().grid();
    mn[0] = g.min.x; mn[1] = g.min.y; mn[2] = g.min.z;
    mx[0] = g.max.x; mx[1] = g.max.y; mx[2] = g.max.z;
    inv[0] = g.invCell.x; inv[1] = g.invCell.y; inv[2] = g.invCell.z;
    size[0] = g.size.x; size[1] = g.size.y; size[2] = g.size.z;
  });
}
int fdmx_get_geometry(void* s, int geo, double* radius, double* invMass, double* I6, double* Ii6,
                      double* origin3) {
  return guarded([&] {
    const fdm_template& t = static_cast<BoundSolver*>(s)->getLSParticle().templateInfo(geo);
    *radius = t.radius;
    *invMass = t.invMass;
    for (int k = 0; k < 6; ++k) {
      I6[k] = t.inertia[k];
      Ii6[k] = t.invInertia[k];
    }
    for (int k = 0; k < 3; ++k) origin3[k] = t.origin[k];
  });
}
// Accumulated stage times in ms (stage timing on); names joined by '\n'.
int fdmx_stage_times(void* s, char* names, int namesCap, double* ms, long long* calls, int cap) {
  int count = 0;
  guarded([&] {
    std::string joined;
    auto* S = static_cast<BoundSolver*>(s);
    const auto& c = S->stageCalls();
    for (const auto& kv : S->stageMilliseconds()) {
      if (count < cap) {
        ms[count] = kv.second;
        auto it = c.find(kv.first);
        calls[count] = it == c.end() ? 0 : it->second;
      }
      joined += kv.first;
      joined += '\n';
      ++count;
    }
    std::strncpy(names, joined.c_str(), (size_t)namesCap - 1);
    names[namesCap - 1] = '\0';
  });
  return count;
}

}  // extern "C"


// OBJ meshes (OBJLoader::loadOBJMesh) and mesh level sets
// (LevelSetObject::TriangleMeshParticle).
namespace {
struct ObjMesh {
  std::vector<double3> v;
  std::vector<int3> f;
};
}  // namespace
extern "C" {
void* fdmx_obj_load(const char* path, int verbose) {
  auto* m = new ObjMesh;
  if (!OBJLoader::loadOBJMesh(path, m->v, m->f, 0, nullptr, verbose != 0)) {
    delete m;
    g_error = std::string("cannot load OBJ mesh ") + path;
    return nullptr;
  }
  return m;
}
long long fdmx_obj_num_vertices(void* h) { return (long long)static_cast<ObjMesh*>(h)->v.size(); }
long long fdmx_obj_num_faces(void* h) { return (long long)static_cast<ObjMesh*>(h)->f.size(); }
void fdmx_obj_get(void* h, double* v, int* f) {
  const ObjMesh& m = *static_cast<ObjMesh*>(h);
  for (size_t i = 0; i < m.v.size(); ++i) {
    v[3 * i] = m.v[i].x;
    v[3 * i + 1] = m.v[i].y;
    v[3 * i + 2] = m.v[i].z;
  }
  for (size_t i = 0; i < m.f.size(); ++i) {
    f[3 * i] = m.f[i].x;
    f[3 * i + 1] = m.f[i].y;
    f[3 * i + 2] = m.f[i].z;
  }
}
void fdmx_obj_free(void* h) { delete static_cast<ObjMesh*>(h); }

void* fdmx_mesh_create(const double* v, int nv, const int* f, int nf, int positiveInside) {
  try {
    auto* m = new LevelSetObject::TriangleMeshParticle(d3v(v, nv), i3v(f, nf));
    m->setLevelSetPositiveInside(positiveInside != 0);
    return m;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void fdmx_mesh_destroy(void* h) { delete static_cast<LevelSetObject::TriangleMeshParticle*>(h); }
int fdmx_mesh_is_valid(void* h) { return static_cast<LevelSetObject::TriangleMeshParticle*>(h)->isValid() ? 1 : 0; }
int fdmx_mesh_build_grid(void* h, double spacing, int padding) {
  return guarded([&] { static_cast<LevelSetObject::TriangleMeshParticle*>(h)->buildGrid(spacing, padding); });
}
int fdmx_mesh_build_grid_by_resolution(void* h, int nodesPerDiameter, int padding) {
  return guarded([&] {
    static_cast<LevelSetObject::TriangleMeshParticle*>(h)->buildGridByResolution(nodesPerDiameter, padding);
  });
}
// origin3, size3, spacing of the built grid (size 0 when none is built)
void fdmx_mesh_grid_info(void* h, double* origin3, int* size3, double* spacing) {
  const auto& g = static_cast<LevelSetObject::TriangleMeshParticle*>(h)->gridInfo();
  origin3[0] = g.gridOrigin.x;
  origin3[1] = g.gridOrigin.y;
  origin3[2] = g.gridOrigin.z;
  const bool built = !g.gridNodeSignedDistance.empty();
  size3[0] = built ? g.gridNodeSize.x : 0;
  size3[1] = built ? g.gridNodeSize.y : 0;
  size3[2] = built ? g.gridNodeSize.z : 0;
  *spacing = g.gridNodeSpacing;
}
void fdmx_mesh_grid_values(void* h, double* out) {
  const auto& g = static_cast<LevelSetObject::TriangleMeshParticle*>(h)->gridInfo();
  std::memcpy(out, g.gridNodeSignedDistance.data(), g.gridNodeSignedDistance.size() * sizeof(double));
}
int fdmx_mesh_evaluate(void* h, const double* points, int n, double* values, double* grads) {
  return guarded([&] {
    static_cast<LevelSetObject::TriangleMeshParticle*>(h)->evaluatePoints(points, n, values, grads);
  });
}
int fdmx_mesh_output_grid_vtu(void* h, const char* prefix) {
  return guarded([&] { static_cast<LevelSetObject::TriangleMeshParticle*>(h)->outputGridVTU(prefix); });
}
}  // extern "C"
