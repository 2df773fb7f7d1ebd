// lkc_tune.cpp -- the launch-shape autotuner.
//
// Some kernels of the path come in forms that give bit-identical results: the integer
// contractions split differently over blocks (exact in any order), the ordered double sums
// unchanged.  The At_T_YnWt_S_Zn_ww contraction runs fused (each block walks every w step)
// or split over its w steps and n parts; a full-layout Gram runs as one block per output
// tile or split over the 1920-blocks.  A cost model picks the form by default.  With tuning
// on, the first launch of each shape class runs every valid form on the caller's stream,
// times each twice with HIP events (forward then reverse order, the faster of the two
// kept), and stores the fastest for the class; later launches of the class take it at
// once.  A shape class is the kernel, each size rounded to half an octave, and the small
// parameters (covariates, digit count) exactly.
//
// Modes: 0 off (the cost model), 1 on (measure new classes), 2 frozen (use what is stored,
// the cost model elsewhere), 3 as 1 and checking that every form wrote the same bits (the
// property the choice rests on).  The initial mode comes from LKC_TUNE (0-3, default 1).
// The store is one per process, shared by every handle and thread; it can be written to a
// text file and read back, so a later process starts tuned.
#include "lkc_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

namespace lkc {

// counts the 8-byte words where a and b differ
__global__ void k_tune_cmp(const unsigned long long* a, const unsigned long long* b, long long n, unsigned long long* bad)
{
  unsigned long long c = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    c += a[i] != b[i];
  if (c) atomicAdd(bad, c);
}

namespace {
struct Entry { int choice; float best_ms; float model_ms; };
std::mutex g_mu;
std::map<TuneKey, Entry> g_store;
int g_mode = -1;
long long g_measured = 0;
double g_saved_ms = 0.0;
long long g_verify_compared = 0, g_verify_mismatched = 0;

int mode_locked()
{
  if (g_mode < 0) {
    const char* e = getenv("LKC_TUNE");
    g_mode = 1;
    if (e && e[0] >= '0' && e[0] <= '3' && e[1] == 0) g_mode = e[0] - '0';
  }
  return g_mode;
}
}  // namespace

int tune_mode()
{
  std::lock_guard<std::mutex> lk(g_mu);
  return mode_locked();
}

void tune_set_mode(int m)
{
  std::lock_guard<std::mutex> lk(g_mu);
  g_mode = m;
}

int tune_bucket(long long x)
{
  if (x <= 1) return 0;
  return (int)std::lround(2.0 * std::log2((double)x));
}

bool tune_get(const TuneKey& k, int* choice)
{
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_store.find(k);
  if (it == g_store.end()) return false;
  *choice = it->second.choice;
  return true;
}

void tune_put(const TuneKey& k, int choice, float best_ms, float model_ms)
{
  std::lock_guard<std::mutex> lk(g_mu);
  g_store[k] = Entry{ choice, best_ms, model_ms };
  g_measured++;
  if (model_ms > best_ms) g_saved_ms += model_ms - best_ms;
}

int tune_measure(hipStream_t st, const int* cand, int n, int model_idx, void (*launch)(void*, int), void* ctx,
                 float* best_ms, float* model_ms, const void* out, size_t out_bytes)
{
  std::vector<hipEvent_t> ev(4 * (size_t)n);
  for (auto& e : ev) if (hipEventCreate(&e) != hipSuccess) e = nullptr;
  // mode 3: a copy of the first form's output, and a device count of differing words
  const bool verify = tune_mode() == 3 && out && out_bytes >= 8 && out_bytes <= ((size_t)512 << 20);
  void* snap = nullptr;
  unsigned long long* bad = nullptr;
  if (verify && (hipMallocAsync(&snap, out_bytes, st) != hipSuccess ||
                 hipMallocAsync((void**)&bad, sizeof(unsigned long long), st) != hipSuccess ||
                 hipMemsetAsync(bad, 0, sizeof(unsigned long long), st) != hipSuccess)) {
    if (snap) (void)hipFreeAsync(snap, st);
    if (bad) (void)hipFreeAsync(bad, st);
    snap = nullptr; bad = nullptr;
  }
  const long long words = (long long)(out_bytes / 8);
  // two rounds: forward, then reverse order (the first form to run also warms the caches)
  for (int r = 0; r < 2; r++)
    for (int j = 0; j < n; j++) {
      const int i = r == 0 ? j : n - 1 - j;
      hipEvent_t a = ev[(size_t)4 * i + 2 * r], b = ev[(size_t)4 * i + 2 * r + 1];
      if (a) (void)hipEventRecord(a, st);
      launch(ctx, cand[i]);
      if (b) (void)hipEventRecord(b, st);
      if (bad) {
        if (r == 0 && j == 0) (void)hipMemcpyAsync(snap, out, out_bytes, hipMemcpyDeviceToDevice, st);
        else {
          const long long blocks = std::min<long long>((words + 255) / 256, 4096);
          hipLaunchKernelGGL(k_tune_cmp, dim3((unsigned)blocks), dim3(256), 0, st, (const unsigned long long*)out,
                             (const unsigned long long*)snap, words, bad);
        }
      }
    }
  int best = model_idx;
  float bms = 1e30f;
  std::vector<float> t(n, 1e30f);
  for (int i = 0; i < n; i++)
    for (int r = 0; r < 2; r++) {
      hipEvent_t a = ev[(size_t)4 * i + 2 * r], b = ev[(size_t)4 * i + 2 * r + 1];
      float ms = 0;
      if (a && b && hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess)
        t[i] = std::min(t[i], ms);
    }
  for (int i = 0; i < n; i++) if (t[i] < bms) { bms = t[i]; best = i; }
  for (auto& e : ev) if (e) (void)hipEventDestroy(e);
  if (bad) {
    unsigned long long h = 0;
    const bool ok = hipMemcpyAsync(&h, bad, sizeof h, hipMemcpyDeviceToHost, st) == hipSuccess &&
                    hipStreamSynchronize(st) == hipSuccess;
    (void)hipFreeAsync(snap, st);
    (void)hipFreeAsync(bad, st);
    std::lock_guard<std::mutex> lk(g_mu);
    g_verify_compared += (long long)(2 * n - 1);
    g_verify_mismatched += ok ? (long long)h : 1;
  }
  *best_ms = bms;
  *model_ms = t[model_idx];
  return best;
}

void tune_verify_stats(long long* compared, long long* mismatched)
{
  std::lock_guard<std::mutex> lk(g_mu);
  *compared = g_verify_compared;
  *mismatched = g_verify_mismatched;
}

int tune_entries()
{
  std::lock_guard<std::mutex> lk(g_mu);
  return (int)g_store.size();
}

void tune_stats(long long* measured, double* saved_ms)
{
  std::lock_guard<std::mutex> lk(g_mu);
  *measured = g_measured;
  *saved_ms = g_saved_ms;
}

void tune_clear()
{
  std::lock_guard<std::mutex> lk(g_mu);
  g_store.clear();
  g_measured = 0;
  g_saved_ms = 0.0;
  g_verify_compared = 0;
  g_verify_mismatched = 0;
}

// text: one class per line, "lkc-tune op a b c d e choice best_ms model_ms"
int tune_save(const char* path)
{
  std::lock_guard<std::mutex> lk(g_mu);
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  for (const auto& kv : g_store)
    fprintf(f, "lkc-tune %d %d %d %d %d %d %d %.6g %.6g\n", kv.first.op, kv.first.a, kv.first.b, kv.first.c,
            kv.first.d, kv.first.e, kv.second.choice, kv.second.best_ms, kv.second.model_ms);
  const int bad = ferror(f);
  fclose(f);
  return bad ? -1 : (int)g_store.size();
}

int tune_load(const char* path)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  int n = 0;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    TuneKey k;
    Entry e;
    if (sscanf(line, "lkc-tune %d %d %d %d %d %d %d %f %f", &k.op, &k.a, &k.b, &k.c, &k.d, &k.e, &e.choice,
               &e.best_ms, &e.model_ms) == 9) {
      g_store[k] = e;
      n++;
    }
  }
  fclose(f);
  return n;
}

}  // namespace lkc
