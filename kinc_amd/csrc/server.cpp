/*
 * server.cpp — continuous-batching query server (include/kinc_server.h).
 *
 * One scheduler thread owns two kinc_sim handles with the expression matrix
 * resident, each with its own stream.  Queries queue in arrival order; a
 * query is admitted when its device result buffer (K and correlations of
 * all its pairs) fits in the HBM budget, which is a first-fit arena carved
 * out of one allocation made at start-up (no allocation on the serving
 * path); a query that does not fit does not hold back later ones that do,
 * until it has waited longer than the aging limit.  Each batch takes whole
 * 32,768-pair work blocks from the admitted queries, fewest remaining blocks
 * first, runs as one persistent-form call with per-block segment starts,
 * and its outputs are scattered into the queries' result buffers.  A batch
 * is one work block (the scheduling quantum) while more than one query has
 * work or waits, and up to batch_pairs pairs when one query runs alone.  The
 * two handles take batches in turn, so a batch starts on the CUs the
 * previous one's last pairs leave idle, and the host packs the next batch
 * while the device runs.  A query whose blocks are all done is copied back
 * on a separate stream and filtered into hits (Similarity::process
 * thresholds, src/core/similarity.cpp:104-123).
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
   Clock::time_point t_first;   /* launch of its first block */
   bool started = false;
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
   int handle = 0;              /* engine handle and stream the batch runs on */
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
   kinc_sim* sim[2] = {};            /* two handles, one per in-flight batch */
   hipStream_t stream[2] = {};       /* batches of each handle */
   hipStream_t copy = nullptr;       /* result copy-back */
   hipEvent_t t0 = nullptr;          /* device time origin of the busy record */
   double busy_until_ms = 0.0;       /* end of the last counted busy interval */
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
   int next_handle = 0;               /* the handle the next batch runs on */
};

/* A query that does not fit in the HBM budget stops holding back later
 * queries after this long in the queue; from then on nothing is admitted
 * ahead of it. */
static const double kAgingSeconds = 5.0;

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
   /* queries with blocks left, fewest remaining blocks first (arrival order
    * among equals); a batch is one block while more than one query has work
    * or another waits for admission, so a new small query waits at most for
    * the batches already in flight */
   std::vector<std::shared_ptr<Query>> cand;
   bool others_waiting = false;
   {
      std::lock_guard<std::mutex> g(s->mu);
      for ( auto& q : s->active )
      {
         if ( q->state == 1 && q->next_block < q->nblocks ) cand.push_back(q);
      }
      others_waiting = !s->pending.empty();
   }
   if ( cand.empty() ) return false;
   std::stable_sort(cand.begin(), cand.end(), [](const std::shared_ptr<Query>& x,
                                                 const std::shared_ptr<Query>& y) {
      const int64_t rx = x->nblocks - x->next_block, ry = y->nblocks - y->next_block;
      return rx != ry ? rx < ry : x->id < y->id;
   });
   const int limit = (cand.size() > 1 || others_waiting) ? KINC_WORK_BLOCK_SIZE : s->cap;
   b.segs.clear();
   b.count = 0;
   for ( auto& q : cand )
   {
      while ( q->next_block < q->nblocks )
      {
         const int64_t lo = q->next_block * KINC_WORK_BLOCK_SIZE;
         const int cnt = (int)std::min<int64_t>(KINC_WORK_BLOCK_SIZE, q->n - lo);
         if ( b.count + cnt > limit ) break;
         memcpy(b.h_pairs + 2 * (size_t)b.count, q->pairs.data() + 2 * lo, sizeof(int32_t) * 2 * (size_t)cnt);
         for ( int i = 0; i < cnt; ++i ) b.h_seg[b.count + i] = b.count;
         b.segs.push_back(Segment{q, q->next_block, b.count, cnt});
         if ( !q->started )
         {
            q->started = true;
            q->t_first = Clock::now();
         }
         b.count += cnt;
         q->next_block++;
      }
      if ( b.count >= limit ) break;
   }
   if ( b.count == 0 ) return false;

   const hipStream_t st = s->stream[b.handle];
   SRV_HIP(hipEventRecord(b.start, st));
   SRV_HIP(hipMemcpyAsync(b.d_pairs, b.h_pairs, sizeof(int32_t) * 2 * (size_t)b.count, hipMemcpyHostToDevice, st));
   SRV_HIP(hipMemcpyAsync(b.d_seg, b.h_seg, sizeof(int32_t) * (size_t)b.count, hipMemcpyHostToDevice, st));
   int rc = kinc_sim_execute_pairs_device(s->sim[b.handle], b.d_pairs, b.d_seg, b.count, b.d_K, nullptr,
                                          b.d_corr, nullptr, st);
   if ( rc != KINC_OK ) throw std::string(kinc_sim_last_error(s->sim[b.handle]));
   for ( auto& sg : b.segs )
   {
      const int64_t lo = sg.block * KINC_WORK_BLOCK_SIZE;
      char* base = s->d_arena + sg.q->off;
      SRV_HIP(hipMemcpyAsync(base + lo, b.d_K + sg.offset, (size_t)sg.count, hipMemcpyDeviceToDevice, st));
      SRV_HIP(hipMemcpyAsync(base + sg.q->corr_off + sizeof(float) * (size_t)lo * s->MC,
                             b.d_corr + (size_t)sg.offset * s->MC, sizeof(float) * (size_t)sg.count * s->MC,
                             hipMemcpyDeviceToDevice, st));
   }
   SRV_HIP(hipEventRecord(b.done, st));
   b.busy = true;
   s->batches++;
   return true;
}

/* A finished batch: count its blocks, start the copy-back of complete queries. */
static void retire_batch(kinc_server* s, Batch& b)
{
   /* device-busy time: the union of the batches' intervals (batches on the
    * two handles overlap), measured from the server's time origin; batches
    * retire in start order */
   float t_start = 0.0f, t_end = 0.0f;
   if ( hipEventElapsedTime(&t_start, s->t0, b.start) == hipSuccess &&
        hipEventElapsedTime(&t_end, s->t0, b.done) == hipSuccess )
   {
      const double from = std::max((double)t_start, s->busy_until_ms);
      std::lock_guard<std::mutex> g(s->mu);   /* kinc_server_stats reads it */
      if ( t_end > from ) s->busy_s += (t_end - from) * 1e-3;
      s->busy_until_ms = std::max(s->busy_until_ms, (double)t_end);
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
         /* retire finished batches in launch order (the older of two busy
          * batches is on next_handle) */
         for ( int i = 0; i < 2; ++i )
         {
            Batch& b = s->batch[(s->next_handle + i) % 2];
            if ( !b.busy ) continue;
            if ( hipEventQuery(b.done) != hipSuccess ) break;
            retire_batch(s, b);
            did = true;
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
         /* admit queued queries in arrival order, each one whose buffer
          * fits; one that does not fit holds back the queries behind it only
          * once it has waited longer than the aging limit */
         {
            std::lock_guard<std::mutex> g(s->mu);
            if ( s->stop ) break;
            const auto now = Clock::now();
            for ( auto it = s->pending.begin(); it != s->pending.end(); )
            {
               auto q = *it;
               size_t off = 0;
               if ( s->arena->alloc(q->bytes, &off) )
               {
                  q->off = off;
                  q->state = 1;
                  s->active.push_back(q);
                  it = s->pending.erase(it);
                  did = true;
                  continue;
               }
               if ( std::chrono::duration<double>(now - q->t_submit).count() > kAgingSeconds ) break;
               ++it;
            }
         }
         /* keep a batch in flight on each handle, the handles in turn */
         for ( int i = 0; i < 2; ++i )
         {
            Batch& b = s->batch[s->next_handle];
            if ( b.busy ) break;
            if ( !launch_batch(s, b) ) break;
            s->next_handle ^= 1;
            did = true;
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
   for ( int i = 0; i < 2; ++i )
   {
      int rc = kinc_sim_create(&p, emx, G, S, device, &s->sim[i]);
      if ( rc != KINC_OK )
      {
         const std::string m = kinc_sim_last_error(nullptr);
         kinc_server_destroy(s.release());
         return server_fail(nullptr, rc, m);
      }
   }
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
   for ( int i = 0; i < 2; ++i ) CR(hipStreamCreateWithFlags(&s->stream[i], hipStreamNonBlocking));
   CR(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
   CR(hipEventCreate(&s->t0));
   CR(hipEventRecord(s->t0, s->stream[0]));
   CR(hipEventSynchronize(s->t0));
   CR(hipMalloc(&s->d_arena, (size_t)hbm_budget_bytes));
   s->budget = (size_t)hbm_budget_bytes & ~(size_t)255;
   s->arena.reset(new Arena((size_t)hbm_budget_bytes));
   for ( int i = 0; i < 2; ++i )
   {
      Batch& b = s->batch[i];
      b.handle = i;
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

extern "C" int kinc_server_timing(kinc_server* s, int64_t id, double* start_wait_s, double* run_s)
{
   if ( !s ) return KINC_E_INVALID;
   std::lock_guard<std::mutex> g(s->mu);
   auto it = s->table.find(id);
   if ( it == s->table.end() ) return KINC_E_INVALID;
   auto q = it->second;
   if ( q->state != 3 ) return KINC_E_INVALID;
   const auto first = q->started ? q->t_first : q->t_done;
   if ( start_wait_s ) *start_wait_s = std::chrono::duration<double>(first - q->t_submit).count();
   if ( run_s ) *run_s = std::chrono::duration<double>(q->t_done - first).count();
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
   for ( int i = 0; i < 2; ++i )
   {
      if ( s->stream[i] ) (void)hipStreamSynchronize(s->stream[i]);
      if ( s->sim[i] ) (void)kinc_sim_synchronize(s->sim[i]);
   }
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
   for ( int i = 0; i < 2; ++i )
   {
      if ( s->stream[i] ) (void)hipStreamDestroy(s->stream[i]);
      kinc_sim_destroy(s->sim[i]);
   }
   if ( s->copy ) (void)hipStreamDestroy(s->copy);
   if ( s->t0 ) (void)hipEventDestroy(s->t0);
   delete s;
}

extern "C" const char* kinc_server_last_error(const kinc_server* s)
{
   return s ? s->err.c_str() : g_server_create_error.c_str();
}
