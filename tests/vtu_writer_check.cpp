// Writes one .vtu file with fundem::vtu::Piece from known arrays (argv[1]),
// for tests/test_vtu.py to decode independently. Array lengths are chosen so
// that the base64 blocks end on every padding case (0, 1 and 2 '=').
#include <cstdint>
#include <string>
#include <vector>

#include "../fundem_amd/include/fundem/VTUWriter.h"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  using namespace fundem;
  const size_t N = 4, M = 2;
  std::vector<float> pts(3 * N), vel(3 * N), q(4 * N), e(N);
  std::vector<int32_t> id(N), conn = {0, 1, 2, 1, 2, 3}, offs = {3, 6};
  std::vector<uint8_t> types(M, 5);
  for (size_t i = 0; i < 3 * N; ++i) {
    pts[i] = 0.25f * float(i) - 1.0f;
    vel[i] = 1e-3f * float(i * i);
  }
  for (size_t i = 0; i < 4 * N; ++i) q[i] = float(i) / 7.0f;
  for (size_t i = 0; i < N; ++i) {
    e[i] = float(i) * 1.5e-9f;
    id[i] = int32_t(100 - 3 * i);
  }
  vtu::makeDirs(argv[1]);
  {
    vtu::Piece f(1.25, 7, N, M);
    f.points(pts);
    f.cells(conn, offs, types);
    f.beginPointData("velocity");
    f.pointArray("particleID", id, 1);
    f.pointArray("velocity", vel, 3);
    f.pointArray("orientation", q, 4);
    f.pointArray("energy", e, 1);
    f.write(vtu::framePath(argv[1], "Mesh", 3));
  }
  {
    vtu::Piece f(0.5, 11, N, 0);
    f.points(pts);
    f.noCells();
    f.beginPointData("");
    f.pointArray("velocity", vel, 3);
    f.write(vtu::framePath(argv[1], "Points", 12));
  }
  {
    vtu::Piece f(0.0, 0, 0, 0);
    f.points(std::vector<float>{});
    f.cells({}, {}, {});
    f.beginPointData("");
    f.pointArray("contactNormal", std::vector<float>{}, 3);
    f.write(vtu::framePath(argv[1], "Empty", 0));
  }
  // a size mismatch is an error, not a malformed file
  try {
    vtu::Piece f(0.0, 0, N + 1, 0);
    f.points(pts);
    return 3;
  } catch (const std::logic_error&) {
  }
  (void)argv[2];
  return 0;
}
