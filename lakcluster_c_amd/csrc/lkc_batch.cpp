// lkc_batch.cpp -- many eliminations on one GPU: permutation-null batches and a serving loop.
//
// Every job is an ordinary elimination (Run) whose inputs are drawn from a device-resident
// pool; jobs share nothing but the pool, so a job's trace is the same alone or in a mix.
// Jobs are admitted in arrival order (jobs arriving together largest first) while their
// memory estimate fits the HBM budget;
// each running job drives its own HIP stream from its own host thread, so the small
// kernels of concurrent jobs overlap on the device.  Arrival times make the same loop a
// server: a job is admitted no earlier than its arrival, latency = end - arrival.
#include "lkc_engine.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>

namespace lkc {

// Device bytes a job of this shape may hold at once: the two double Grams, the bit
// matrices in both orientations, the ww digits of the initial pass and the rest.
double job_bytes(int nA, int nZ, int M, int B)
{
  const double a = nA, z = nZ, m = M, nd = (B + 7) / 8;
  const double grams = 8.0 * a * (a + z);
  const double bits = 2.0 * (a + z) * (double)bsize(M) + m * ((double)bsize(nA) + (double)bsize(nZ));
  const double digits = nd * (std::max(a, z) + 32.0) * 8.0 * (double)std::max(bsize(nA), bsize(nZ));
  const double per_col = 8.0 * m * 40.0 + 8.0 * (a + z) * 40.0;
  return grams + bits + digits + per_col + 256.0 * 1024 * 1024;
}

void run_batch(const Pool& P, const JobIn* jobs, int n, int max_conc, double budget, JobOut* out)
{
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto since = [&](clk::time_point t) { return std::chrono::duration<double>(t - t0).count(); };
  std::mutex mu;
  std::condition_variable cv;
  int next = 0, running = 0;
  double reserved = 0;
  // admission order: by arrival; jobs that arrive together (a batch) largest first, so the
  // last jobs to finish are short ones and the device stays full to the end
  std::vector<int> order(n);
  std::vector<double> work(n);
  for (int i = 0; i < n; i++) {
    order[i] = i;
    int cols = 0;
    if (jobs[i].colmask) { for (int k = 0; k < P.M; k++) cols += (jobs[i].colmask[k >> 3] >> (7 - (k & 7))) & 1; }
    else cols = P.M;
    work[i] = (double)jobs[i].nA * (double)(jobs[i].nA + jobs[i].nZ) * (double)cols;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (jobs[x].arrival_s != jobs[y].arrival_s) return jobs[x].arrival_s < jobs[y].arrival_s;
    return work[x] > work[y];
  });
  auto worker = [&]() {
    for (;;) {
      std::unique_lock<std::mutex> lk(mu);
      if (next >= n) return;
      // in admission order: wait for the head job's arrival, then for room in the budget
      const int j = order[next];
      const auto arrive = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(jobs[j].arrival_s));
      while (next < n && order[next] == j && clk::now() < arrive) cv.wait_until(lk, arrive);
      if (next >= n || order[next] != j) continue;
      const double need = job_bytes(jobs[j].nA, jobs[j].nZ, P.M, jobs[j].cfg.B);
      while (next < n && order[next] == j && running > 0 && reserved + need > budget) cv.wait(lk);
      if (next >= n || order[next] != j) continue;
      next++; reserved += need; running++;
      lk.unlock();
      cv.notify_all();
      JobOut& o = out[j];
      o.arrival_s = jobs[j].arrival_s;
      o.start_s = since(clk::now());
      try {
        Run r;
        r.load_job(P, jobs[j].rowsA, jobs[j].nA, jobs[j].rowsZ, jobs[j].nZ, jobs[j].colmask, jobs[j].cfg);
        r.run(-1);
        o.trace = r.trace; o.xdrop_a = r.xdrop_a; o.xdrop_b = r.xdrop_b;
        o.n_iter = r.n_iter; o.ops = r.ops; o.status = LKC_OK;
      } catch (const Error& e) {
        o.status = e.code; o.err = e.what();
      } catch (const std::exception& e) {
        o.status = LKC_ERR_HIP; o.err = e.what();
      }
      o.end_s = since(clk::now());
      lk.lock();
      reserved -= need; running--;
      lk.unlock();
      cv.notify_all();
    }
  };
  const int nt = std::max(1, std::min(max_conc, n));
  std::vector<std::thread> th;
  for (int i = 0; i < nt; i++) th.emplace_back(worker);
  for (auto& t : th) t.join();
}

}  // namespace lkc
