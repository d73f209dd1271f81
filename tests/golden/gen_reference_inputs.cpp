// Regenerates the seeded inputs of two reference unit tests with the same libstdc++ engine and
// distributions (std::mt19937 + uniform_real_distribution<float>), so the oracle can be checked on
// exactly those inputs:
//   broadphase.bin   tests/arborx/test_broadphase_arborx.cpp:19-40  (seed 12345, 4000 particles,
//                    x, y, z ~ U[0,20), r ~ U[0.3,0.7), drawn per particle in that order)
//   periodicity.bin  tests/kokkos/test_periodicity.cpp:20-61 (seed 23, 600 reals biased to the
//                    faces of [-5,5]^3; only the positions are kept)
// Output: little-endian float32 arrays.  Build: g++ -O2 -o gen gen_reference_inputs.cpp
#include <cstdio>
#include <random>
#include <vector>

int main(int argc, char** argv) {
  const char* dir = argc > 1 ? argv[1] : ".";
  char path[1024];
  {
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> upos(0.f, 20.f);
    std::uniform_real_distribution<float> urad(0.3f, 0.7f);
    std::vector<float> out;
    for (int i = 0; i < 4000; ++i) {
      const float x = upos(rng), y = upos(rng), z = upos(rng), r = urad(rng);
      out.push_back(x); out.push_back(y); out.push_back(z); out.push_back(r);
    }
    std::snprintf(path, sizeof path, "%s/broadphase.bin", dir);
    FILE* f = std::fopen(path, "wb");
    std::fwrite(out.data(), sizeof(float), out.size(), f);
    std::fclose(f);
  }
  {
    std::mt19937 rng(23);
    std::uniform_real_distribution<float> uf(-5.f, 5.f);
    std::uniform_real_distribution<float> uedge(0.f, 1.f);
    std::normal_distribution<float> nq(0.f, 1.f);
    auto coord = [&]() {
      const float r = uedge(rng);
      if (r < 0.25f) return -5.f + uedge(rng) * 0.4f;
      if (r < 0.5f) return 5.f - uedge(rng) * 0.4f;
      return uf(rng);
    };
    std::vector<float> out;
    for (int i = 0; i < 600; ++i) {
      const float x = coord();
      const float y = coord();
      const float z = coord();
      out.push_back(x); out.push_back(y); out.push_back(z);
      for (int k = 0; k < 3; ++k) (void)uf(rng);        // velocity
      for (int k = 0; k < 4; ++k) (void)nq(rng);        // quaternion
      (void)uedge(rng);                                 // scale
    }
    std::snprintf(path, sizeof path, "%s/periodicity.bin", dir);
    FILE* f = std::fopen(path, "wb");
    std::fwrite(out.data(), sizeof(float), out.size(), f);
    std::fclose(f);
  }
  return 0;
}
