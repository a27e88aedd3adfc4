// Host build of the Lanczos solve's dense Ritz-vector routine (specann_amd/csrc/sa_ritz.hpp), for the
// CPU test that checks it against numpy's eigh. dense_ritz_host runs it on one thread;
// dense_ritz_team runs it on `threads` host threads that meet at a barrier wherever the device's
// wavefront meets at a wave barrier, which checks the work split and the barrier placement of the
// device form (64 threads is the device's team size).
#include <barrier>
#include <thread>
#include <vector>

#include "sa_ritz.hpp"

namespace {

struct HostTeam {
  int r, n;
  std::barrier<>* bar;
  int rank() const { return r; }
  int size() const { return n; }
  void sync() const { bar->arrive_and_wait(); }
};

}  // namespace

extern "C" int dense_ritz_host(int m, int k, const double* a, const double* b, const double* ar, double theta,
                               double tiny, double* x) {
  if (m < 1 || m > 128 || k < 0 || k >= m) return -1;
  std::vector<double> M((size_t)m * m);
  std::vector<int> perm(m);
  sa::dense_ritz_vector(sa::SerialTeam{}, m, k, a, b, ar, theta, tiny, M.data(), perm.data(), x);
  return 0;
}

extern "C" int dense_ritz_team(int threads, int m, int k, const double* a, const double* b, const double* ar,
                               double theta, double tiny, double* x) {
  if (threads < 1 || threads > 256 || m < 1 || m > 128 || k < 0 || k >= m) return -1;
  std::vector<double> M((size_t)m * m);
  std::vector<int> perm(m);
  std::barrier<> bar(threads);
  std::vector<std::thread> pool;
  for (int r = 0; r < threads; ++r)
    pool.emplace_back([&, r] {
      sa::dense_ritz_vector(HostTeam{r, threads, &bar}, m, k, a, b, ar, theta, tiny, M.data(), perm.data(), x);
    });
  for (auto& th : pool) th.join();
  return 0;
}
