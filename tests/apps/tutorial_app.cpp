// A FunDEM application written against the C++ host API as the reference's
// tutorials are (apps/tutorial/LSParticleRolling.cpp:42-62): an LSSolver
// subclass whose addExternalForceTorque override launches the tutorials'
// global damping on getLSParticle()'s device arrays with its gridDim() and
// blockDim(), particles added as clusters (LSSolver::addParticleCluster, one
// geometry per cluster, placed by its density), a wall, and solve().
//
// The scene comes from a binary input file (written by the test that runs
// this program); the program writes the particles' initial positions (as
// addParticleCluster placed them), the final state after solve(), and the
// launch shapes and contact copies the getters report.
//
//   tutorial_app <input> <output> <frame directory>
#include <cstdio>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fundem/LSSolver.h"
#include "fundem/globalDamping.h"

class solver : public LSSolver {
 public:
  explicit solver(const std::string& dir) : LSSolver(dir) {}

  double dampingCoefficient = 0.2;

  void addExternalForceTorque(const double time) override {
    (void)time;
    launchAddGlobalDampingForceTorque(getLSParticle().force(), getLSParticle().torque(),
                                      getLSParticle().velocity(),
                                      getLSParticle().angularVelocity(), dampingCoefficient,
                                      getLSParticle().num_device(), getLSParticle().gridDim(),
                                      getLSParticle().blockDim(), 0);
  }
};

namespace {

struct Reader {
  std::ifstream in;
  explicit Reader(const char* path) : in(path, std::ios::binary) {
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  }
  template <class T>
  T get() {
    T v;
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::runtime_error("input truncated");
    return v;
  }
  double3 d3() {
    const double x = get<double>(), y = get<double>(), z = get<double>();
    return make_double3(x, y, z);
  }
  quaternion q() {
    const double a = get<double>(), b = get<double>(), c = get<double>(), d = get<double>();
    return make_quaternion(a, b, c, d);
  }
};

struct Grid {
  double3 origin;
  double spacing;
  int3 size;
  std::vector<double> sdf;
  std::vector<double3> nodes;
  std::vector<int3> tris;
};

Grid readGrid(Reader& r) {
  Grid g;
  g.origin = r.d3();
  g.spacing = r.get<double>();
  const int nx = r.get<int32_t>(), ny = r.get<int32_t>(), nz = r.get<int32_t>();
  g.size = make_int3(nx, ny, nz);
  g.sdf.resize(size_t(nx) * ny * nz);
  for (double& v : g.sdf) v = r.get<double>();
  g.nodes.resize(r.get<int32_t>());
  for (double3& p : g.nodes) p = r.d3();
  g.tris.resize(r.get<int32_t>());
  for (int3& t : g.tris) {
    const int a = r.get<int32_t>(), b = r.get<int32_t>(), c = r.get<int32_t>();
    t = make_int3(a, b, c);
  }
  return g;
}

void put(std::ofstream& out, double v) { out.write(reinterpret_cast<const char*>(&v), sizeof v); }
void put3(std::ofstream& out, double3 v) {
  put(out, v.x);
  put(out, v.y);
  put(out, v.z);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: tutorial_app <input> <output> <frame directory>" << std::endl;
    return 2;
  }
  Reader r(argv[1]);
  solver s(argv[3]);
  s.dampingCoefficient = r.get<double>();
  const double density = r.get<double>();
  const double kn = r.get<double>(), ks = r.get<double>(), mu = r.get<double>(), e = r.get<double>();
  s.addParticleMaterial(kn, ks, mu, e);
  const double wmu = r.get<double>(), we = r.get<double>();
  s.addWallMaterial(wmu, we);
  // one cluster per template: the particles of that template
  const int nTemplates = r.get<int32_t>();
  for (int t = 0; t < nTemplates; ++t) {
    const Grid g = readGrid(r);
    const int n = r.get<int32_t>();
    std::vector<double3> p(n), v(n), w(n);
    std::vector<quaternion> q(n);
    for (int i = 0; i < n; ++i) {
      p[i] = r.d3();
      v[i] = r.d3();
      w[i] = r.d3();
      q[i] = r.q();
    }
    s.addParticleCluster(p, v, w, q, 0, density, g.origin, g.spacing, g.size, g.sdf, g.nodes,
                         g.tris);
  }
  const Grid wall = readGrid(r);
  s.addWallGeometry(wall.origin, wall.spacing, wall.size, wall.sdf, wall.nodes, wall.tris);
  const double3 wx = r.d3();
  s.addWall(wx, make_double3(0., 0., 0.), make_double3(0., 0., 0.), make_quaternion(1., 0., 0., 0.),
            0, 0);
  const double3 mn = r.d3(), mx = r.d3(), gravity = r.d3();
  const double dt = r.get<double>(), maxTime = r.get<double>();
  const int numFrame = r.get<int32_t>();

  std::ofstream out(argv[2], std::ios::binary);
  const auto& P = s.getLSParticle();
  const size_t n = P.num();
  for (size_t i = 0; i < n; ++i) put3(out, P.positionHostRef()[i]);

  s.solve(mn, mx, gravity, dt, maxTime, size_t(numFrame), argc, argv);
  // solve() ended on an output frame, which downloaded the state and the
  // contacts' host copies (LSParticle / SolidInteraction HostRef getters)
  for (size_t i = 0; i < n; ++i) {
    put3(out, P.positionHostRef()[i]);
    put3(out, P.velocityHostRef()[i]);
    put3(out, P.angularVelocityHostRef()[i]);
    const quaternion q = P.orientationHostRef()[i];
    put(out, q.q0);
    put(out, q.q1);
    put(out, q.q2);
    put(out, q.q3);
    put3(out, P.forceHostRef()[i]);
    put3(out, P.torqueHostRef()[i]);
  }
  auto& pw = s.getLSParticleWallInteraction();
  auto& pp = s.getLSParticleInteraction();
  put(out, double(P.gridDim()));
  put(out, double(P.blockDim()));
  for (auto* si : {&pp, &pw}) {
    const std::vector<int> master = si->masterIDHostRef(), slave = si->slaveIDHostRef();
    const std::vector<double3> point = si->contactPointHostRef(), normal = si->contactNormalHostRef();
    put(out, double(si->numPair_device()));
    put(out, double(si->pairGridDim()));
    put(out, double(si->pairBlockDim()));
    for (size_t k = 0; k < master.size(); ++k) {
      put(out, double(master[k]));
      put(out, double(slave[k]));
      put3(out, point[k]);
      put3(out, normal[k]);
    }
  }
  if (!out) {
    std::cerr << "write failed" << std::endl;
    return 1;
  }
  return 0;
}
