/*
 * server.cpp — continuous-batching query server (include/kinc_server.h).
 *
 * One scheduler thread owns a kinc_sim handle with the expression matrix
 * resident.  Queries queue in arrival order; a query is admitted when its
 * device result buffer (K and correlations of all its pairs) fits in the
 * HBM budget, which is a first-fit arena carved out of one allocation made
 * at start-up (no allocation on the serving path).  Each batch takes whole
 * 32,768-pair work blocks round-robin from the admitted queries, runs as one
 * persistent-form call with per-block segment starts, and its outputs are
 * scattered into the queries' result buffers.  Two batches are in flight, so
 * the host packs batch b+1 while the device runs batch b.  A query whose
 * blocks are all done is copied back on a separate stream and filtered into
 * hits (Similarity::process thresholds, src/core/similarity.cpp:104-123).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/kinc_server.h"
#include "../../include/kinc_sim.h"

namespace {

using Clock = std::chrono::steady_clock;

thread_local std::string g_server_create_error;

/* First-fit allocator over [0, size) in 256-byte units. */
class Arena
{
public:
   explicit Arena(size_t size) { free_[0] = size; }

   bool alloc(size_t bytes, size_t* off)
   {
      bytes = (bytes + 255) & ~(size_t)255;
      for ( auto it = free_.begin(); it != free_.end(); ++it )
      {
         if ( it->second >= bytes )
         {
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if ( rest ) free_[at] = rest;
            return true;
         }
      }
      return false;
   }

   void release(size_t off, size_t bytes)
   {
      bytes = (bytes + 255) & ~(size_t)255;
      auto it = free_.emplace(off, bytes).first;
      auto next = std::next(it);
      if ( next != free_.end() && it->first + it->second == next->first )
      {
         it->second += next->second;
         free_.erase(next);
      }
      if ( it != free_.begin() )
      {
         auto prev = std::prev(it);
         if ( prev->first + prev->second == it->first )
         {
            prev->second += it->second;
            free_.erase(it);
         }
      }
   }

private:
   std::map<size_t, size_t> free_;
};

struct Query
{
   int64_t id = 0;
   float min_abs = 0.5f;
   std::vector<int32_t> pairs;  /* 2 per pair */
   int64_t n = 0;
   int64_t nblocks = 0;
   int64_t next_block = 0;
   int64_t blocks_done = 0;
   size_t off = 0;              /* arena offset: K[n], then corr[n][MC] (256-aligned) */
   size_t bytes = 0;
   size_t corr_off = 0;
   int state = 0;               /* 0 queued, 1 admitted, 2 copying, 3 done, -1 failed */
   Clock::time_point t_submit, t_done;
   std::vector<int8_t> hK;
   std::vector<float> hcorr;
   hipEvent_t copied = nullptr;
   std::vector<kinc_hit> hits;
};

struct Segment
{
   std::shared_ptr<Query> q;
   int64_t block;
   int offset;                  /* first pair in the batch */
   int count;
};

struct Batch
{
   int32_t* h_pairs = nullptr;  /* pinned */
   int32_t* h_seg = nullptr;    /* pinned */
   int32_t* d_pairs = nullptr;
   int32_t* d_seg = nullptr;
   int8_t* d_K = nullptr;
   float* d_corr = nullptr;
   hipEvent_t start = nullptr;
   hipEvent_t done = nullptr;
   bool busy = false;
   int count = 0;
   std::vector<Segment> segs;
};

} // namespace

struct kinc_server
{
   kinc_sim_params p;
   int device = 0;
   int G = 0;
   int S = 0;
   int MC = 0;
   int cap = 0;
   kinc_sim* sim = nullptr;
   hipStream_t stream = nullptr;     /* batches */
   hipStream_t copy = nullptr;       /* result copy-back */
   char* d_arena = nullptr;
   size_t budget = 0;
   std::unique_ptr<Arena> arena;
   Batch batch[2];

   std::mutex mu;
   std::condition_variable cv;       /* new work or a finished query */
   std::deque<std::shared_ptr<Query>> pending;
   std::vector<std::shared_ptr<Query>> active;
   std::unordered_map<int64_t, std::shared_ptr<Query>> table;
   int64_t next_id = 1;
   bool stop = false;
   bool dead = false;                /* the scheduler stopped on a device error */
   std::thread worker;
   std::string err;

   std::atomic<int64_t> pairs_done{0};
   std::atomic<int64_t> batches{0};
   double busy_s = 0.0;
   size_t rr = 0;                     /* round-robin position */
};

static int server_fail(kinc_server* s, int code, const std::string& m)
{
   if ( s )
   {
      std::lock_guard<std::mutex> g(s->mu);
      s->err = m;
   }
   else
   {
      g_server_create_error = m;
   }
   return code;
}

/* The pair list of a query, as documented in kinc_server.h. */
static void build_pairs(const int32_t* genes, int n_genes, int G, Query& q)
{
   std::vector<int32_t> pos((size_t)G, -1);
   std::vector<int32_t> uniq;
   for ( int i = 0; i < n_genes; ++i )
   {
      const int g = genes[i];
      if ( pos[(size_t)g] < 0 )
      {
         pos[(size_t)g] = (int32_t)uniq.size();
         uniq.push_back(g);
      }
   }
   q.pairs.clear();
   for ( size_t i = 0; i < uniq.size(); ++i )
   {
      const int a = uniq[i];
      for ( int g = 0; g < G; ++g )
      {
         if ( g == a ) continue;
         if ( pos[(size_t)g] >= 0 && (size_t)pos[(size_t)g] < i ) continue;
         q.pairs.push_back(std::max(a, g));
         q.pairs.push_back(std::min(a, g));
      }
   }
   q.n = (int64_t)q.pairs.size() / 2;
   q.nblocks = (q.n + KINC_WORK_BLOCK_SIZE - 1) / KINC_WORK_BLOCK_SIZE;
}

/* Similarity::process thresholds per pair; kept clusters renumbered. */
static void filter_hits(kinc_server* s, Query& q)
{
   q.hits.clear();
   const int MC = s->MC;
   for ( int64_t i = 0; i < q.n; ++i )
   {
      const int K = q.hK[(size_t)i];
      int c = 0;
      for ( int k = 0; k < K; ++k )
      {
         const float r = q.hcorr[(size_t)i * MC + k];
         if ( !std::isnan(r) && q.min_abs <= std::fabs(r) && std::fabs(r) <= s->p.max_corr )
         {
            kinc_hit h;
            h.x = q.pairs[(size_t)(2 * i)];
            h.y = q.pairs[(size_t)(2 * i + 1)];
            h.cluster = c++;
            h.corr = r;
            q.hits.push_back(h);
         }
      }
   }
   q.hK.clear();
   q.hK.shrink_to_fit();
   q.hcorr.clear();
   q.hcorr.shrink_to_fit();
}

#define SRV_HIP(call)                                                               \
   do                                                                               \
   {                                                                                \
      hipError_t e_ = (call);                                                       \
      if ( e_ != hipSuccess ) throw std::string(#call) + ": " + hipGetErrorString(e_); \
   } while ( 0 )

static bool launch_batch(kinc_server* s, Batch& b)
{
   /* pack whole blocks of admitted queries, round-robin */
   std::vector<std::shared_ptr<Query>> cand;
   {
      std::lock_guard<std::mutex> g(s->mu);
      for ( auto& q : s->active )
      {
         if ( q->state == 1 && q->next_block < q->nblocks ) cand.push_back(q);
      }
   }
   if ( cand.empty() ) return false;
   b.segs.clear();
   b.count = 0;
   bool progress = true;
   size_t start = s->rr % cand.size();
   while ( progress )
   {
      progress = false;
      for ( size_t t = 0; t < cand.size(); ++t )
      {
         auto& q = cand[(start + t) % cand.size()];
         if ( q->next_block >= q->nblocks ) continue;
         const int64_t lo = q->next_block * KINC_WORK_BLOCK_SIZE;
         const int cnt = (int)std::min<int64_t>(KINC_WORK_BLOCK_SIZE, q->n - lo);
         if ( b.count + cnt > s->cap ) continue;
         memcpy(b.h_pairs + 2 * (size_t)b.count, q->pairs.data() + 2 * lo, sizeof(int32_t) * 2 * (size_t)cnt);
         for ( int i = 0; i < cnt; ++i ) b.h_seg[b.count + i] = b.count;
         b.segs.push_back(Segment{q, q->next_block, b.count, cnt});
         b.count += cnt;
         q->next_block++;
         progress = true;
      }
   }
   s->rr++;
   if ( b.count == 0 ) return false;

   SRV_HIP(hipEventRecord(b.start, s->stream));
   SRV_HIP(hipMemcpyAsync(b.d_pairs, b.h_pairs, sizeof(int32_t) * 2 * (size_t)b.count, hipMemcpyHostToDevice, s->stream));
   SRV_HIP(hipMemcpyAsync(b.d_seg, b.h_seg, sizeof(int32_t) * (size_t)b.count, hipMemcpyHostToDevice, s->stream));
   int rc = kinc_sim_execute_pairs_device(s->sim, b.d_pairs, b.d_seg, b.count, b.d_K, nullptr, b.d_corr,
                                          nullptr, s->stream);
   if ( rc != KINC_OK ) throw std::string(kinc_sim_last_error(s->sim));
   for ( auto& sg : b.segs )
   {
      const int64_t lo = sg.block * KINC_WORK_BLOCK_SIZE;
      char* base = s->d_arena + sg.q->off;
      SRV_HIP(hipMemcpyAsync(base + lo, b.d_K + sg.offset, (size_t)sg.count, hipMemcpyDeviceToDevice, s->stream));
      SRV_HIP(hipMemcpyAsync(base + sg.q->corr_off + sizeof(float) * (size_t)lo * s->MC,
                             b.d_corr + (size_t)sg.offset * s->MC, sizeof(float) * (size_t)sg.count * s->MC,
                             hipMemcpyDeviceToDevice, s->stream));
   }
   SRV_HIP(hipEventRecord(b.done, s->stream));
   b.busy = true;
   s->batches++;
   return true;
}

/* A finished batch: count its blocks, start the copy-back of complete queries. */
static void retire_batch(kinc_server* s, Batch& b)
{
   float ms = 0.0f;
   if ( hipEventElapsedTime(&ms, b.start, b.done) == hipSuccess )
   {
      std::lock_guard<std::mutex> g(s->mu);   /* kinc_server_stats reads it */
      s->busy_s += ms * 1e-3;
   }
   s->pairs_done += b.count;
   for ( auto& sg : b.segs )
   {
      auto& q = sg.q;
      q->blocks_done++;
      if ( q->blocks_done == q->nblocks )
      {
         q->hK.resize((size_t)q->n);
         q->hcorr.resize((size_t)q->n * s->MC);
         const char* base = s->d_arena + q->off;
         SRV_HIP(hipMemcpyAsync(q->hK.data(), base, (size_t)q->n, hipMemcpyDeviceToHost, s->copy));
         SRV_HIP(hipMemcpyAsync(q->hcorr.data(), base + q->corr_off, sizeof(float) * (size_t)q->n * s->MC,
                                hipMemcpyDeviceToHost, s->copy));
         SRV_HIP(hipEventCreateWithFlags(&q->copied, hipEventDisableTiming));
         SRV_HIP(hipEventRecord(q->copied, s->copy));
         std::lock_guard<std::mutex> g(s->mu);
         q->state = 2;
      }
   }
   b.segs.clear();
   b.busy = false;
}

static void scheduler(kinc_server* s)
{
   try
   {
      SRV_HIP(hipSetDevice(s->device));
      for ( ;; )
      {
         bool did = false;
         /* retire finished batches */
         for ( auto& b : s->batch )
         {
            if ( b.busy && hipEventQuery(b.done) == hipSuccess )
            {
               retire_batch(s, b);
               did = true;
            }
         }
         /* finish copied queries */
         std::vector<std::shared_ptr<Query>> copying;
         {
            std::lock_guard<std::mutex> g(s->mu);
            for ( auto& q : s->active ) if ( q->state == 2 ) copying.push_back(q);
         }
         for ( auto& q : copying )
         {
            if ( hipEventQuery(q->copied) != hipSuccess ) continue;
            (void)hipEventDestroy(q->copied);
            q->copied = nullptr;
            filter_hits(s, *q);
            std::lock_guard<std::mutex> g(s->mu);
            s->arena->release(q->off, q->bytes);
            q->t_done = Clock::now();
            q->state = 3;
            s->active.erase(std::remove(s->active.begin(), s->active.end(), q), s->active.end());
            did = true;
            s->cv.notify_all();
         }
         /* admit queued queries in arrival order while their buffers fit */
         {
            std::lock_guard<std::mutex> g(s->mu);
            if ( s->stop ) break;
            while ( !s->pending.empty() )
            {
               auto& q = s->pending.front();
               size_t off = 0;
               if ( !s->arena->alloc(q->bytes, &off) ) break;
               q->off = off;
               q->state = 1;
               s->active.push_back(q);
               s->pending.pop_front();
               did = true;
            }
         }
         /* keep two batches in flight */
         for ( auto& b : s->batch )
         {
            if ( !b.busy && launch_batch(s, b) ) did = true;
         }
         if ( !did )
         {
            std::unique_lock<std::mutex> g(s->mu);
            const bool idle = s->pending.empty() && s->active.empty();
            s->cv.wait_for(g, idle ? std::chrono::milliseconds(20) : std::chrono::microseconds(200));
            if ( s->stop ) break;
         }
      }
   }
   catch ( const std::string& e )
   {
      std::lock_guard<std::mutex> g(s->mu);
      s->err = e;
      s->dead = true;
      for ( auto& q : s->active ) q->state = -1;
      for ( auto& q : s->pending ) q->state = -1;
      s->cv.notify_all();
   }
}

extern "C" int kinc_server_create(const kinc_sim_params* params, const float* emx, int32_t G, int32_t S,
                                  int32_t device, int64_t hbm_budget_bytes, int32_t batch_pairs,
                                  kinc_server** out)
{
   if ( !params || !emx || !out ) return server_fail(nullptr, KINC_E_INVALID, "null argument");
   if ( batch_pairs < KINC_WORK_BLOCK_SIZE ) return server_fail(nullptr, KINC_E_INVALID, "batch_pairs must be >= 32768");
   if ( hbm_budget_bytes < (int64_t)1 << 20 ) return server_fail(nullptr, KINC_E_INVALID, "hbm_budget_bytes must be >= 1 MiB");
   *out = nullptr;
   kinc_sim_params p = *params;
   p.batch_capacity = batch_pairs;
   p.form = KINC_FORM_PERSISTENT;
   std::unique_ptr<kinc_server> s(new kinc_server());
   s->p = p;
   s->device = device;
   s->G = G;
   s->S = S;
   s->MC = p.max_clusters;
   s->cap = batch_pairs;
   int rc = kinc_sim_create(&p, emx, G, S, device, &s->sim);
   if ( rc != KINC_OK ) return server_fail(nullptr, rc, kinc_sim_last_error(nullptr));
   auto cleanup_fail = [&](const std::string& m) {
      g_server_create_error = m;
      kinc_server_destroy(s.release());
      return KINC_E_DEVICE;
   };
#define CR(call)                                                                  \
   do                                                                             \
   {                                                                              \
      hipError_t e_ = (call);                                                     \
      if ( e_ != hipSuccess ) return cleanup_fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
   } while ( 0 )
   CR(hipSetDevice(device));
   CR(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
   CR(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
   CR(hipMalloc(&s->d_arena, (size_t)hbm_budget_bytes));
   s->budget = (size_t)hbm_budget_bytes & ~(size_t)255;
   s->arena.reset(new Arena((size_t)hbm_budget_bytes));
   for ( auto& b : s->batch )
   {
      CR(hipHostMalloc(&b.h_pairs, sizeof(int32_t) * 2 * (size_t)batch_pairs, 0));
      CR(hipHostMalloc(&b.h_seg, sizeof(int32_t) * (size_t)batch_pairs, 0));
      CR(hipMalloc(&b.d_pairs, sizeof(int32_t) * 2 * (size_t)batch_pairs));
      CR(hipMalloc(&b.d_seg, sizeof(int32_t) * (size_t)batch_pairs));
      CR(hipMalloc(&b.d_K, (size_t)batch_pairs));
      CR(hipMalloc(&b.d_corr, sizeof(float) * (size_t)batch_pairs * s->MC));
      CR(hipEventCreate(&b.start));
      CR(hipEventCreate(&b.done));
   }
#undef CR
   kinc_server* raw = s.release();
   raw->worker = std::thread(scheduler, raw);
   *out = raw;
   return KINC_OK;
}

extern "C" int kinc_server_submit(kinc_server* s, const int32_t* genes, int32_t n_genes,
                                  float min_abs_corr, int64_t* id)
{
   if ( !s || !genes || !id || n_genes < 1 ) return server_fail(s, KINC_E_INVALID, "invalid query");
   for ( int i = 0; i < n_genes; ++i )
   {
      if ( genes[i] < 0 || genes[i] >= s->G ) return server_fail(s, KINC_E_INVALID, "query gene out of range");
   }
   auto q = std::make_shared<Query>();
   q->min_abs = min_abs_corr;
   q->t_submit = Clock::now();
   build_pairs(genes, n_genes, s->G, *q);
   const size_t kbytes = ((size_t)q->n + 255) & ~(size_t)255;
   q->corr_off = kbytes;
   q->bytes = kbytes + sizeof(float) * (size_t)q->n * s->MC;
   if ( q->bytes > s->budget )
   {
      return server_fail(s, KINC_E_NOMEM, "query needs " + std::to_string(q->bytes)
                                         + " result bytes, more than the HBM budget");
   }
   std::lock_guard<std::mutex> g(s->mu);
   if ( s->dead ) return KINC_E_DEVICE;
   q->id = s->next_id++;
   s->table[q->id] = q;
   if ( q->n == 0 )
   {
      q->state = 3;
      q->t_done = Clock::now();
   }
   else
   {
      s->pending.push_back(q);
   }
   *id = q->id;
   s->cv.notify_all();
   return KINC_OK;
}

extern "C" int kinc_server_wait(kinc_server* s, int64_t id, double timeout_s)
{
   if ( !s ) return KINC_E_INVALID;
   std::unique_lock<std::mutex> g(s->mu);
   auto it = s->table.find(id);
   if ( it == s->table.end() ) return KINC_E_INVALID;
   auto q = it->second;
   const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
   s->cv.wait_until(g, until, [&] { return q->state == 3 || q->state < 0; });
   if ( q->state < 0 ) return KINC_E_DEVICE;
   return q->state == 3 ? 1 : 0;
}

extern "C" int kinc_server_result(kinc_server* s, int64_t id, kinc_hit* hits, int64_t cap,
                                  int64_t* n_hits, double* latency_s, int64_t* n_pairs)
{
   if ( !s ) return KINC_E_INVALID;
   std::lock_guard<std::mutex> g(s->mu);
   auto it = s->table.find(id);
   if ( it == s->table.end() ) return KINC_E_INVALID;
   auto q = it->second;
   if ( q->state != 3 ) return KINC_E_INVALID;
   if ( n_hits ) *n_hits = (int64_t)q->hits.size();
   if ( latency_s ) *latency_s = std::chrono::duration<double>(q->t_done - q->t_submit).count();
   if ( n_pairs ) *n_pairs = q->n;
   if ( hits )
   {
      const int64_t n = std::min<int64_t>(cap, (int64_t)q->hits.size());
      if ( n > 0 ) memcpy(hits, q->hits.data(), sizeof(kinc_hit) * (size_t)n);
      if ( n == (int64_t)q->hits.size() ) s->table.erase(it);
   }
   return KINC_OK;
}

extern "C" int kinc_server_stats(kinc_server* s, int64_t* pairs_done, int64_t* batches, double* busy_s)
{
   if ( !s ) return KINC_E_INVALID;
   if ( pairs_done ) *pairs_done = s->pairs_done.load();
   if ( batches ) *batches = s->batches.load();
   std::lock_guard<std::mutex> g(s->mu);
   if ( busy_s ) *busy_s = s->busy_s;
   return KINC_OK;
}

extern "C" void kinc_server_destroy(kinc_server* s)
{
   if ( !s ) return;
   {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
      s->cv.notify_all();
   }
   if ( s->worker.joinable() ) s->worker.join();
   (void)hipSetDevice(s->device);
   if ( s->stream ) (void)hipStreamSynchronize(s->stream);
   if ( s->copy ) (void)hipStreamSynchronize(s->copy);
   for ( auto& b : s->batch )
   {
      (void)hipHostFree(b.h_pairs);
      (void)hipHostFree(b.h_seg);
      (void)hipFree(b.d_pairs);
      (void)hipFree(b.d_seg);
      (void)hipFree(b.d_K);
      (void)hipFree(b.d_corr);
      if ( b.start ) (void)hipEventDestroy(b.start);
      if ( b.done ) (void)hipEventDestroy(b.done);
   }
   for ( auto& kv : s->table )
   {
      if ( kv.second->copied ) (void)hipEventDestroy(kv.second->copied);
   }
   (void)hipFree(s->d_arena);
   if ( s->stream ) (void)hipStreamDestroy(s->stream);
   if ( s->copy ) (void)hipStreamDestroy(s->copy);
   kinc_sim_destroy(s->sim);
   delete s;
}

extern "C" const char* kinc_server_last_error(const kinc_server* s)
{
   return s ? s->err.c_str() : g_server_create_error.c_str();
}
