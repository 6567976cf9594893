// rand_stream_check.cpp -- host check of jssp::RandStream's process-wide form
// and of jssp::ProcessRandGuard
// (beamsearch_jssp_gpu_amd/csrc/rand_stream.hpp): it must draw exactly the
// values the next rand() calls would return, and leave the process-wide
// generator where those calls would, even when something else in the process
// draws from or reseeds the generator while the stream is alive.
// TEST INFRASTRUCTURE: built by beamsearch_jssp_gpu_amd.build, run by
// tests/test_gpu_drop_in.py.  Exit status 0 = every check passed.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../beamsearch_jssp_gpu_amd/csrc/rand_stream.hpp"

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

// Values rand() returns after srand(seed), skipping `skip`.
static std::vector<int> expected(unsigned seed, int skip, int count) {
  std::srand(seed);
  for (int i = 0; i < skip; ++i) std::rand();
  std::vector<int> v(count);
  for (int& x : v) x = std::rand();
  return v;
}

static void run(unsigned seed, int skip, int count, bool disturb) {
  const std::vector<int> want = expected(seed, skip, count + 3);
  std::srand(seed);
  for (int i = 0; i < skip; ++i) std::rand();
  {
    jssp::RandStream rs(true, 0);
    for (int i = 0; i < count; ++i) {
      if (disturb && i % 7 == 3) {  // another library reseeds and draws meanwhile
        std::srand(424242u + i);
        std::rand();
      }
      CHECK(rs.next() == want[i]);
    }
  }
  for (int i = 0; i < 3; ++i) CHECK(std::rand() == want[count + i]);
}

int main() {
  for (unsigned seed : {0u, 1u, 12345u, 2147483648u, 4294967295u})
    for (int skip : {0, 1, 30, 31, 1000})
      for (bool disturb : {false, true}) run(seed, skip, 2000, disturb);
  // a private seeded stream never touches the process-wide one
  std::srand(77);
  const int a = std::rand();
  std::srand(77);
  {
    jssp::RandStream rs(false, 5);
    const std::vector<int> want = expected(5, 0, 10);
    std::srand(77);
    for (int i = 0; i < 10; ++i) CHECK(rs.next() == want[i]);
  }
  CHECK(std::rand() == a);
  // ProcessRandGuard: the generator resumes where it was, whatever happened
  // inside (reseeds, draws, another state block installed)
  for (unsigned seed : {3u, 12345u}) {
    const std::vector<int> want = expected(seed, 10, 20);
    std::srand(seed);
    for (int i = 0; i < 10; ++i) std::rand();
    {
      jssp::ProcessRandGuard keep;
      std::srand(1);
      for (int i = 0; i < 100; ++i) std::rand();
      static char other[64];
      initstate(7, other, sizeof other);
      std::rand();
    }
    for (int i = 0; i < 20; ++i) CHECK(std::rand() == want[i]);
  }
  // a caller-installed larger state (TYPE_4, 256 bytes)
  static char big[256];
  initstate(99, big, sizeof big);
  std::vector<int> want(50);
  for (int& x : want) x = std::rand();
  initstate(99, big, sizeof big);
  for (int i = 0; i < 10; ++i) std::rand();
  {
    jssp::RandStream rs(true, 0);
    for (int i = 10; i < 40; ++i) CHECK(rs.next() == want[i]);
  }
  for (int i = 40; i < 50; ++i) CHECK(std::rand() == want[i]);
  if (fails) return 1;
  std::printf("ok\n");
  return 0;
}
