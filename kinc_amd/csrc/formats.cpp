/*
 * formats.cpp — EMX reader/writer and the CCM/CMX writer of the Similarity
 * analytic (include/kinc_formats.h).
 *
 * The writer is Similarity::process (src/core/similarity.cpp:66-127) over
 * Pairwise::Matrix::write / Pair::write (src/core/pairwise_matrix.cpp:122-151,
 * 251-273): per pair in index order, clusters whose correlation is not NaN
 * and satisfies min_corr <= |r| <= max_corr are kept and renumbered from 0;
 * a CCM record is written only when K > 1, a CMX record when K > 0; each
 * record's indent (x(x-1)/2 + y)*64 + cluster must exceed the previous one.
 *
 * The text importer/exporter is ImportExpressionMatrix::process
 * (src/core/importexpressionmatrix.cpp:20-170) and
 * ExportExpressionMatrix::process (src/core/exportexpressionmatrix.cpp:20-99),
 * with the lines of a file parsed by several threads.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/kinc_formats.h"
#include "../../include/kinc_sim.h"

/* Pairs of a process() call per encoding thread, at least (thread start-up
 * costs about as much as encoding this many single-cluster pairs). */
#ifndef KINC_WRITER_PAIRS_PER_THREAD
#define KINC_WRITER_PAIRS_PER_THREAD 2048
#endif

static_assert(sizeof(float) == 4, "float32");

namespace {

thread_local std::string g_err;

int ferr(int code, const std::string& m)
{
   g_err = m;
   return code;
}

const char kMagic[8] = {'K', 'I', 'N', 'C', 'A', 'M', 'D', '1'};
const uint64_t kPayloadOffset = 64;

/* little-endian field writers (the host is x86-64 or another LE target) */
void put32(std::vector<uint8_t>& b, uint32_t v)
{
   for ( int i = 0; i < 4; ++i ) b.push_back((uint8_t)(v >> (8 * i)));
}

void put64(std::vector<uint8_t>& b, uint64_t v)
{
   for ( int i = 0; i < 8; ++i ) b.push_back((uint8_t)(v >> (8 * i)));
}

uint32_t get32(const uint8_t* p)
{
   return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

uint64_t get64(const uint8_t* p)
{
   return (uint64_t)get32(p) | ((uint64_t)get32(p + 4) << 32);
}

std::vector<uint8_t> envelope(uint32_t type, uint64_t payload_size, uint64_t meta_size)
{
   std::vector<uint8_t> b(kMagic, kMagic + 8);
   put32(b, type);
   put32(b, 0);
   put64(b, kPayloadOffset);
   put64(b, payload_size);
   put64(b, kPayloadOffset + payload_size);
   put64(b, meta_size);
   b.resize(kPayloadOffset, 0);
   return b;
}

struct EnvelopeInfo
{
   uint32_t type;
   uint64_t payload_off, payload_size, meta_off, meta_size;
};

int read_envelope(FILE* f, EnvelopeInfo& e, const char* path)
{
   uint8_t h[64];
   if ( fread(h, 1, 64, f) != 64 ) return ferr(KINC_E_INVALID, std::string(path) + ": short file");
   if ( memcmp(h, kMagic, 8) != 0 ) return ferr(KINC_E_INVALID, std::string(path) + ": not a KINCAMD1 file");
   e.type = get32(h + 8);
   e.payload_off = get64(h + 16);
   e.payload_size = get64(h + 24);
   e.meta_off = get64(h + 32);
   e.meta_size = get64(h + 40);
   return KINC_OK;
}

/* Insert `"key": value,` at the front of a JSON object string. */
std::string json_with(const std::string& obj, const std::string& key, const std::string& value)
{
   std::string o = obj.empty() ? "{}" : obj;
   size_t p = o.find('{');
   if ( p == std::string::npos ) o = "{}", p = 0;
   std::string rest = o.substr(p + 1);
   size_t q = rest.find_first_not_of(" \t\r\n");
   bool empty = q != std::string::npos && rest[q] == '}';
   return o.substr(0, p + 1) + "\"" + key + "\": " + value + (empty ? "" : ", ") + rest;
}

/* One pairwise matrix file (Pairwise::Matrix). */
struct PairwiseFile
{
   FILE* f = nullptr;
   std::string path;
   uint32_t type = 0;
   int32_t G = 0;
   int32_t max_clusters = 0;
   int32_t data_size = 0;
   int64_t pair_size = 0;
   int64_t cluster_size = 0;
   std::vector<uint8_t> sub;   /* sub-header bytes */
   int64_t last_write = -1;
   std::string meta;
   std::vector<uint8_t> rec;

   std::vector<uint8_t> header() const
   {
      std::vector<uint8_t> b;
      put32(b, (uint32_t)G);
      put32(b, (uint32_t)max_clusters);
      put32(b, (uint32_t)data_size);
      put64(b, (uint64_t)pair_size);
      put64(b, (uint64_t)cluster_size);
      b.push_back((uint8_t)(sub.size() & 0xff));
      b.push_back((uint8_t)(sub.size() >> 8));
      b.insert(b.end(), sub.begin(), sub.end());
      return b;
   }

   int open(const char* p)
   {
      path = p;
      f = fopen(p, "wb");
      if ( !f ) return ferr(KINC_E_INVALID, path + ": cannot open for writing");
      std::vector<uint8_t> e = envelope(type, 0, 0);
      std::vector<uint8_t> h = header();
      if ( fwrite(e.data(), 1, e.size(), f) != e.size() || fwrite(h.data(), 1, h.size(), f) != h.size() )
      {
         return ferr(KINC_E_INVALID, path + ": write failed");
      }
      return KINC_OK;
   }

   /* Matrix::write (pairwise_matrix.cpp:122-151) + the record's data */
   int write(int32_t x, int32_t y, int cluster, const uint8_t* data)
   {
      const int64_t indent = ((int64_t)x * (x - 1) / 2 + y) * KINC_MAX_CLUSTER_SIZE + cluster;
      if ( indent <= last_write )
      {
         return ferr(KINC_E_INVALID, "Attempting to write indent " + std::to_string(indent)
                                     + " when last written is " + std::to_string(last_write));
      }
      rec.clear();
      put32(rec, (uint32_t)x);
      put32(rec, (uint32_t)y);
      rec.push_back((uint8_t)(int8_t)cluster);
      rec.insert(rec.end(), data, data + data_size);
      if ( fwrite(rec.data(), 1, rec.size(), f) != rec.size() ) return ferr(KINC_E_INVALID, path + ": write failed");
      ++cluster_size;
      last_write = indent;
      return KINC_OK;
   }

   /* Records encoded elsewhere, in increasing indent order from first to
    * last: Matrix::write's order check on the first, then one fwrite. */
   int append(const std::string& bytes, int64_t records, int64_t first_indent, int64_t last_indent)
   {
      if ( records == 0 ) return KINC_OK;
      if ( first_indent <= last_write )
      {
         return ferr(KINC_E_INVALID, "Attempting to write indent " + std::to_string(first_indent)
                                     + " when last written is " + std::to_string(last_write));
      }
      if ( fwrite(bytes.data(), 1, bytes.size(), f) != bytes.size() ) return ferr(KINC_E_INVALID, path + ": write failed");
      cluster_size += records;
      last_write = last_indent;
      return KINC_OK;
   }

   /* Matrix::finish (pairwise_matrix.cpp:51-57): final counts, metadata */
   int close()
   {
      if ( !f ) return KINC_OK;
      std::vector<uint8_t> h = header();
      const uint64_t payload = h.size() + (uint64_t)cluster_size * (uint64_t)(9 + data_size);
      int rc = KINC_OK;
      if ( fwrite(meta.data(), 1, meta.size(), f) != meta.size() ) rc = ferr(KINC_E_INVALID, path + ": write failed");
      std::vector<uint8_t> e = envelope(type, payload, meta.size());
      if ( fseek(f, 0, SEEK_SET) != 0 || fwrite(e.data(), 1, e.size(), f) != e.size()
           || fwrite(h.data(), 1, h.size(), f) != h.size() )
      {
         rc = ferr(KINC_E_INVALID, path + ": header rewrite failed");
      }
      if ( fclose(f) != 0 ) rc = ferr(KINC_E_INVALID, path + ": close failed");
      f = nullptr;
      return rc;
   }
};

} // namespace

namespace {

/* The records of a run of pairs, encoded into memory. */
struct Encoded
{
   std::string ccm, cmx;
   int64_t ccm_recs = 0, ccm_pairs = 0, cmx_recs = 0, cmx_pairs = 0;
   int64_t ccm_first = -1, ccm_last = -1, cmx_first = -1, cmx_last = -1;
   int64_t bad = -1;        /* first pair that cannot be written */
   std::string why;
   /* empty again, keeping the buffers' capacity */
   void reset()
   {
      ccm.clear();
      cmx.clear();
      ccm_recs = ccm_pairs = cmx_recs = cmx_pairs = 0;
      ccm_first = ccm_last = cmx_first = cmx_last = -1;
      bad = -1;
      why.clear();
   }
};

/* A fixed set of worker threads that run the tasks of one call at a time;
 * the caller runs task 0 itself and waits for the rest. */
class WorkerPool
{
public:
   explicit WorkerPool(int workers)
   {
      for ( int t = 1; t <= workers; ++t ) th_.emplace_back([this, t] { loop(t); });
   }
   ~WorkerPool()
   {
      {
         std::lock_guard<std::mutex> g(m_);
         stop_ = true;
      }
      cv_.notify_all();
      for ( auto& t : th_ ) t.join();
   }
   int size() const { return (int)th_.size() + 1; }
   /* f(t) for t in [0, tasks), tasks <= size() */
   void run(int tasks, const std::function<void(int)>& f)
   {
      {
         std::lock_guard<std::mutex> g(m_);
         job_ = &f;
         tasks_ = tasks;
         pending_ = tasks - 1;
         ++gen_;
      }
      cv_.notify_all();
      f(0);
      std::unique_lock<std::mutex> g(m_);
      done_.wait(g, [this] { return pending_ == 0; });
      job_ = nullptr;
   }

private:
   void loop(int t)
   {
      uint64_t seen = 0;
      for ( ;; )
      {
         const std::function<void(int)>* job = nullptr;
         {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [&] { return stop_ || gen_ != seen; });
            if ( stop_ ) return;
            seen = gen_;
            if ( t < tasks_ ) job = job_;
         }
         if ( job )
         {
            (*job)(t);
            std::lock_guard<std::mutex> g(m_);
            if ( --pending_ == 0 ) done_.notify_one();
         }
      }
   }
   std::vector<std::thread> th_;
   std::mutex m_;
   std::condition_variable cv_, done_;
   const std::function<void(int)>* job_ = nullptr;
   int tasks_ = 0;
   int pending_ = 0;
   uint64_t gen_ = 0;
   bool stop_ = false;
};

} // namespace

struct kinc_similarity_writer
{
   PairwiseFile ccm;
   PairwiseFile cmx;
   int32_t S = 0;
   int32_t max_clusters = 0;
   float min_corr = 0.5f;
   float max_corr = 1.0f;
   /* encoding threads and their buffers, kept across process() calls */
   std::unique_ptr<WorkerPool> pool;
   std::vector<Encoded> parts;
};

extern "C" const char* kinc_formats_last_error(void)
{
   return g_err.c_str();
}

extern "C" int kinc_emx_write(const char* path, const float* data, int32_t G, int32_t S,
                              const char* metadata_json)
{
   if ( !path || !data || G < 1 || S < 1 ) return ferr(KINC_E_INVALID, "invalid EMX arguments");
   FILE* f = fopen(path, "wb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open for writing");
   const std::string meta = metadata_json ? metadata_json : "{}";
   const uint64_t payload = 8 + (uint64_t)G * (uint64_t)S * 4;
   std::vector<uint8_t> e = envelope(KINC_FILE_EMX, payload, meta.size());
   std::vector<uint8_t> h;
   put32(h, (uint32_t)G);
   put32(h, (uint32_t)S);
   bool ok = fwrite(e.data(), 1, e.size(), f) == e.size() && fwrite(h.data(), 1, h.size(), f) == h.size()
          && fwrite(data, 4, (size_t)G * (size_t)S, f) == (size_t)G * (size_t)S
          && fwrite(meta.data(), 1, meta.size(), f) == meta.size();
   ok = (fclose(f) == 0) && ok;
   return ok ? KINC_OK : ferr(KINC_E_INVALID, std::string(path) + ": write failed");
}

extern "C" int kinc_emx_info(const char* path, int32_t* G, int32_t* S)
{
   FILE* f = fopen(path, "rb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open");
   EnvelopeInfo e;
   int rc = read_envelope(f, e, path);
   uint8_t h[8];
   if ( rc == KINC_OK && e.type != KINC_FILE_EMX ) rc = ferr(KINC_E_INVALID, std::string(path) + ": not an EMX file");
   if ( rc == KINC_OK && (fseek(f, (long)e.payload_off, SEEK_SET) != 0 || fread(h, 1, 8, f) != 8) )
   {
      rc = ferr(KINC_E_INVALID, std::string(path) + ": short file");
   }
   fclose(f);
   if ( rc ) return rc;
   *G = (int32_t)get32(h);
   *S = (int32_t)get32(h + 4);
   return KINC_OK;
}

extern "C" int kinc_emx_read(const char* path, float* data)
{
   int32_t G = 0, S = 0;
   int rc = kinc_emx_info(path, &G, &S);
   if ( rc ) return rc;
   FILE* f = fopen(path, "rb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open");
   EnvelopeInfo e;
   rc = read_envelope(f, e, path);
   const size_t n = (size_t)G * (size_t)S;
   if ( rc == KINC_OK && (fseek(f, (long)(e.payload_off + 8), SEEK_SET) != 0 || fread(data, 4, n, f) != n) )
   {
      rc = ferr(KINC_E_INVALID, std::string(path) + ": short file");
   }
   fclose(f);
   return rc;
}

extern "C" int64_t kinc_file_metadata(const char* path, char* buf, int64_t cap)
{
   FILE* f = fopen(path, "rb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open");
   EnvelopeInfo e;
   int rc = read_envelope(f, e, path);
   if ( rc ) { fclose(f); return rc; }
   if ( buf && cap > 0 )
   {
      const size_t n = (size_t)std::min<uint64_t>(e.meta_size, (uint64_t)cap);
      if ( fseek(f, (long)e.meta_off, SEEK_SET) != 0 || fread(buf, 1, n, f) != n )
      {
         fclose(f);
         return ferr(KINC_E_INVALID, std::string(path) + ": short file");
      }
   }
   fclose(f);
   return (int64_t)e.meta_size;
}

extern "C" int kinc_similarity_writer_open(const char* ccm_path, const char* cmx_path,
                                           int32_t G, int32_t S, int32_t max_clusters,
                                           int32_t corr_method, float min_corr, float max_corr,
                                           const char* metadata_json, kinc_similarity_writer** out)
{
   if ( !ccm_path || !cmx_path || !out || G < 2 || S < 1 ) return ferr(KINC_E_INVALID, "invalid writer arguments");
   /* Matrix::initialize (pairwise_matrix.cpp:72-110) */
   if ( max_clusters < 1 || max_clusters > KINC_MAX_CLUSTER_SIZE )
   {
      return ferr(KINC_E_INVALID, "The max cluster size cannot be larger than MAX_CLUSTER_SIZE.");
   }
   kinc_similarity_writer* w = new kinc_similarity_writer();
   w->S = S;
   w->max_clusters = max_clusters;
   w->min_corr = min_corr;
   w->max_corr = max_corr;
   const std::string meta = metadata_json ? metadata_json : "{}";

   w->ccm.type = KINC_FILE_CCM;
   w->ccm.G = G;
   w->ccm.max_clusters = max_clusters;
   w->ccm.data_size = (S + 1) / 2;
   put32(w->ccm.sub, (uint32_t)S);
   w->ccm.meta = meta;

   w->cmx.type = KINC_FILE_CMX;
   w->cmx.G = G;
   w->cmx.max_clusters = max_clusters;
   w->cmx.data_size = 4;
   w->cmx.sub.push_back(1);
   w->cmx.meta = json_with(meta, "correlations",
                           corr_method == KINC_CORR_SPEARMAN ? "[\"spearman\"]" : "[\"pearson\"]");

   int rc = w->ccm.open(ccm_path);
   if ( rc == KINC_OK ) rc = w->cmx.open(cmx_path);
   if ( rc )
   {
      if ( w->ccm.f ) fclose(w->ccm.f);
      if ( w->cmx.f ) fclose(w->cmx.f);
      delete w;
      return rc;
   }
   *out = w;
   return KINC_OK;
}

namespace {

/* CCM nibble of a sample label for cluster k (Similarity::process,
 * similarity.cpp:92-99): 1/0 for membership of a clustered sample, the
 * negated marker (9 NaN, 8 post-outlier, 7 pre-outlier, 6 below minimum)
 * otherwise, in 4 bits.  Two samples per byte, so the table maps a pair of
 * labels (little-endian uint16) straight to the mask byte. */
struct NibbleTable
{
   static constexpr int kTables = 8;   /* clusters 0..7; higher ones take the direct path */
   uint8_t t[kTables][65536];
   static uint8_t nib(int k, int8_t lab)
   {
      const int v = (lab >= 0) ? (k == lab ? 1 : 0) : -(int)lab;
      return (uint8_t)(v & 0x0f);
   }
   NibbleTable()
   {
      for ( int k = 0; k < kTables; ++k )
      {
         for ( int u = 0; u < 65536; ++u )
         {
            t[k][u] = (uint8_t)(nib(k, (int8_t)(u & 0xff)) | (nib(k, (int8_t)(u >> 8)) << 4));
         }
      }
   }
};

const NibbleTable& nibble_table()
{
   static const NibbleTable table;
   return table;
}

inline void put_header(std::string& o, int32_t x, int32_t y, int c)
{
   char h[9];
   for ( int i = 0; i < 4; ++i ) h[i] = (char)((uint32_t)x >> (8 * i));
   for ( int i = 0; i < 4; ++i ) h[4 + i] = (char)((uint32_t)y >> (8 * i));
   h[8] = (char)(int8_t)c;
   o.append(h, 9);
}

} // namespace

/*
 * Similarity::process (similarity.cpp:66-127) for pairs [start, start+count).
 * Threads encode contiguous runs of pairs into memory; the runs are then
 * appended to the files in pair order, so the files are byte-identical to a
 * serial pass.  A pair that cannot be written stops the call after every
 * record before it has been written, as the serial loop would.
 */
extern "C" int kinc_similarity_writer_process(kinc_similarity_writer* w, int64_t start, int64_t count,
                                              const int8_t* K, const int8_t* labels,
                                              const float* corr, int32_t corr_stride)
{
   if ( !w || !K || !corr || count < 0 ) return ferr(KINC_E_INVALID, "invalid process arguments");
   if ( count == 0 ) return KINC_OK;
   const int S = w->S;
   const int nb = (S + 1) / 2;
   const NibbleTable& lut = nibble_table();
   const float lo = w->min_corr, hi = w->max_corr;
   const int MC = w->max_clusters;

   auto encode = [&](int64_t q0, int64_t q1, Encoded& e) {
      int32_t x, y;
      kinc_pair_index(start + q0, &x, &y);
      e.ccm.reserve((size_t)(q1 - q0) * 16);
      e.cmx.reserve((size_t)(q1 - q0) * 20);
      for ( int64_t q = q0; q < q1; ++q )
      {
         const int k_sel = K[q];
         const float* r = corr + q * corr_stride;
         if ( k_sel == 1 )
         {
            /* a run of single-cluster pairs: no CCM record, at most one CMX
             * record each, written straight into the buffer */
            size_t at = e.cmx.size();
            e.cmx.resize(at + (size_t)(q1 - q) * 13);
            char* o = &e.cmx[0];
            int64_t first = e.cmx_first, last = e.cmx_last, recs = 0;
            for ( ; q < q1 && K[q] == 1; ++q )
            {
               /* branch-free: the record is always written and kept by
                * advancing past it (a NaN fails both comparisons) */
               const float c = corr[q * corr_stride];
               const float a = std::fabs(c);
               const bool keep = (lo <= a) & (a <= hi);
               const uint32_t ux = (uint32_t)x, uy = (uint32_t)y;
               memcpy(o + at, &ux, 4);      /* little-endian host */
               memcpy(o + at + 4, &uy, 4);
               o[at + 8] = 0;
               memcpy(o + at + 9, &c, 4);
               at += keep ? 13 : 0;
               const int64_t indent = ((int64_t)x * (x - 1) / 2 + y) * KINC_MAX_CLUSTER_SIZE;
               first = (keep & (first < 0)) ? indent : first;
               last = keep ? indent : last;
               recs += keep;
               if ( ++y >= x )
               {
                  y = 0;
                  ++x;
               }
            }
            e.cmx_first = first;
            e.cmx_last = last;
            e.cmx_recs += recs;
            e.cmx_pairs += recs;
            e.cmx.resize(at);
            --q;
            continue;
         }
         if ( k_sel > MC )
         {
            e.bad = q;
            e.why = "Cannot write pair with cluster size exceeding the max";
            return;
         }
         if ( k_sel > 1 && !labels )
         {
            e.bad = q;
            e.why = "labels are required when K > 1";
            return;
         }
         const int64_t base_indent = ((int64_t)x * (x - 1) / 2 + y) * KINC_MAX_CLUSTER_SIZE;
         int n = 0;
         for ( int k = 0; k < k_sel; ++k )
         {
            const float c = r[k];
            if ( std::isnan(c) || !(lo <= std::fabs(c) && std::fabs(c) <= hi) ) continue;
            if ( k_sel > 1 )
            {
               const uint8_t* lab = (const uint8_t*)(labels + q * S);
               put_header(e.ccm, x, y, n);
               const size_t at = e.ccm.size();
               e.ccm.resize(at + (size_t)nb);
               uint8_t* m = (uint8_t*)&e.ccm[at];
               const int pairs2 = S / 2;
               if ( k < NibbleTable::kTables )
               {
                  const uint8_t* tk = lut.t[k];
                  for ( int i = 0; i < pairs2; ++i ) m[i] = tk[lab[2 * i] | ((unsigned)lab[2 * i + 1] << 8)];
               }
               else
               {
                  for ( int i = 0; i < pairs2; ++i )
                  {
                     m[i] = (uint8_t)(NibbleTable::nib(k, (int8_t)lab[2 * i])
                                      | (NibbleTable::nib(k, (int8_t)lab[2 * i + 1]) << 4));
                  }
               }
               if ( S & 1 ) m[pairs2] = NibbleTable::nib(k, (int8_t)lab[S - 1]);
               if ( e.ccm_first < 0 ) e.ccm_first = base_indent + n;
               e.ccm_last = base_indent + n;
               ++e.ccm_recs;
            }
            put_header(e.cmx, x, y, n);
            e.cmx.append((const char*)&c, 4);
            if ( e.cmx_first < 0 ) e.cmx_first = base_indent + n;
            e.cmx_last = base_indent + n;
            ++e.cmx_recs;
            ++n;
         }
         if ( n > 0 )
         {
            if ( k_sel > 1 ) ++e.ccm_pairs;
            ++e.cmx_pairs;
         }
         /* Pairwise::Index::operator++ */
         if ( ++y >= x )
         {
            y = 0;
            ++x;
         }
      }
   };

   if ( !w->pool )
   {
      int hw = (int)std::thread::hardware_concurrency();
      w->pool.reset(new WorkerPool(std::max(1, std::min(hw, 16)) - 1));
      w->parts.resize((size_t)w->pool->size());
   }
   int T = w->pool->size();
   T = (int)std::min<int64_t>(T, std::max<int64_t>(1, count / KINC_WRITER_PAIRS_PER_THREAD));
   std::vector<Encoded>& parts = w->parts;
   for ( int t = 0; t < T; ++t ) parts[(size_t)t].reset();
   if ( T == 1 ) encode(0, count, parts[0]);
   else
   {
      w->pool->run(T, [&](int t) { encode(count * t / T, count * (t + 1) / T, parts[(size_t)t]); });
   }
   for ( int t = 0; t < T; ++t )
   {
      const Encoded& e = parts[(size_t)t];
      int rc = w->ccm.append(e.ccm, e.ccm_recs, e.ccm_first, e.ccm_last);
      if ( rc == KINC_OK ) rc = w->cmx.append(e.cmx, e.cmx_recs, e.cmx_first, e.cmx_last);
      if ( rc ) return rc;
      w->ccm.pair_size += e.ccm_pairs;
      w->cmx.pair_size += e.cmx_pairs;
      if ( e.bad >= 0 ) return ferr(KINC_E_INVALID, e.why);
   }
   return KINC_OK;
}

extern "C" int kinc_similarity_writer_close(kinc_similarity_writer* w)
{
   if ( !w ) return KINC_E_INVALID;
   int rc1 = w->ccm.close();
   int rc2 = w->cmx.close();
   delete w;
   return rc1 ? rc1 : rc2;
}

/* ------------------------------------------------ expression matrix text */

namespace {

/* QRegExp("\\s+") word separators for ASCII text (QChar::isSpace). */
inline bool is_space(char c)
{
   return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

struct Word
{
   const char* b;
   const char* e;
   std::string str() const { return std::string(b, e); }
   bool equals(const std::string& t) const
   {
      return (size_t)(e - b) == t.size() && memcmp(b, t.data(), t.size()) == 0;
   }
};

/* line.splitRef(QRegExp("\\s+"), QString::SkipEmptyParts) */
void split_words(const char* b, const char* e, std::vector<Word>& out)
{
   out.clear();
   const char* p = b;
   while ( p < e )
   {
      while ( p < e && is_space(*p) ) ++p;
      if ( p == e ) break;
      const char* q = p;
      while ( q < e && !is_space(*q) ) ++q;
      out.push_back({p, q});
      p = q;
   }
}

const double kPow10[23] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11,
                           1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

/*
 * A decimal number in C-locale syntax ([+-] digits [. digits] [(e|E) [+-]
 * digits], at least one mantissa digit) or nan / inf / +inf / -inf, read to
 * the nearest double.  Numbers whose digits fit in 2^53 with a decimal
 * exponent of at most 22 are exact products/quotients of two doubles
 * (Clinger's fast path); anything else goes to strtod, which rounds
 * correctly too.  Returns false for anything else (hex, grouping, text).
 */
bool parse_double(const char* b, const char* e, double& out)
{
   const size_t n = (size_t)(e - b);
   if ( n == 0 ) return false;
   if ( (n == 3 && memcmp(b, "nan", 3) == 0) )
   {
      out = NAN;
      return true;
   }
   if ( (n == 3 && memcmp(b, "inf", 3) == 0) || (n == 4 && memcmp(b, "+inf", 4) == 0) )
   {
      out = INFINITY;
      return true;
   }
   if ( n == 4 && memcmp(b, "-inf", 4) == 0 )
   {
      out = -INFINITY;
      return true;
   }
   const char* p = b;
   bool neg = false;
   if ( *p == '+' || *p == '-' )
   {
      neg = *p == '-';
      ++p;
   }
   uint64_t mant = 0;
   int digits = 0;     /* significant digits accumulated in mant */
   int dropped = 0;    /* integer digits beyond 19 (scale up) */
   int frac = 0;       /* fraction digits taken into mant */
   bool any = false;
   bool exact = true;
   while ( p < e && *p >= '0' && *p <= '9' )
   {
      any = true;
      if ( digits < 19 )
      {
         if ( mant != 0 || *p != '0' ) ++digits;
         mant = mant * 10 + (uint64_t)(*p - '0');
      }
      else
      {
         ++dropped;
         if ( *p != '0' ) exact = false;
      }
      ++p;
   }
   if ( p < e && *p == '.' )
   {
      ++p;
      while ( p < e && *p >= '0' && *p <= '9' )
      {
         any = true;
         if ( digits < 19 )
         {
            if ( mant != 0 || *p != '0' ) ++digits;
            mant = mant * 10 + (uint64_t)(*p - '0');
            ++frac;
         }
         else if ( *p != '0' )
         {
            exact = false;
         }
         ++p;
      }
   }
   if ( !any ) return false;
   int ex = 0;
   if ( p < e && (*p == 'e' || *p == 'E') )
   {
      ++p;
      bool eneg = false;
      if ( p < e && (*p == '+' || *p == '-') )
      {
         eneg = *p == '-';
         ++p;
      }
      if ( p == e ) return false;
      while ( p < e && *p >= '0' && *p <= '9' )
      {
         if ( ex < 100000 ) ex = ex * 10 + (*p - '0');
         ++p;
      }
      if ( eneg ) ex = -ex;
   }
   if ( p != e ) return false;
   const int e10 = ex + dropped - frac;
   if ( exact && mant <= (1ull << 53) && e10 >= -22 && e10 <= 22 )
   {
      double v = (double)mant;
      v = (e10 >= 0) ? v * kPow10[e10] : v / kPow10[-e10];
      out = neg ? -v : v;
      return true;
   }
   std::string tmp(b, e);
   out = strtod(tmp.c_str(), nullptr);
   return true;
}

const char* kTransformNames[4] = {"none", "natural logarithm", "logarithm base 2", "logarithm base 10"};

void json_string(std::string& o, const std::string& s)
{
   o.push_back('"');
   for ( unsigned char c : s )
   {
      if ( c == '"' || c == '\\' )
      {
         o.push_back('\\');
         o.push_back((char)c);
      }
      else if ( c < 0x20 )
      {
         char u[8];
         snprintf(u, sizeof(u), "\\u%04x", c);
         o += u;
      }
      else
      {
         o.push_back((char)c);
      }
   }
   o.push_back('"');
}

void json_array(std::string& o, const std::vector<std::string>& v)
{
   o.push_back('[');
   for ( size_t i = 0; i < v.size(); ++i )
   {
      if ( i ) o.push_back(',');
      json_string(o, v[i]);
   }
   o.push_back(']');
}

/* Append code point cp as UTF-8. */
void put_utf8(std::string& o, uint32_t cp)
{
   if ( cp < 0x80 ) o.push_back((char)cp);
   else if ( cp < 0x800 )
   {
      o.push_back((char)(0xc0 | (cp >> 6)));
      o.push_back((char)(0x80 | (cp & 0x3f)));
   }
   else if ( cp < 0x10000 )
   {
      o.push_back((char)(0xe0 | (cp >> 12)));
      o.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
      o.push_back((char)(0x80 | (cp & 0x3f)));
   }
   else
   {
      o.push_back((char)(0xf0 | (cp >> 18)));
      o.push_back((char)(0x80 | ((cp >> 12) & 0x3f)));
      o.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
      o.push_back((char)(0x80 | (cp & 0x3f)));
   }
}

/* A JSON string starting at s[i] == '"'; on success i is past its end. */
bool json_read_string(const std::string& s, size_t& i, std::string& out)
{
   out.clear();
   if ( i >= s.size() || s[i] != '"' ) return false;
   ++i;
   while ( i < s.size() )
   {
      char c = s[i++];
      if ( c == '"' ) return true;
      if ( c != '\\' )
      {
         out.push_back(c);
         continue;
      }
      if ( i >= s.size() ) return false;
      char x = s[i++];
      switch ( x )
      {
      case '"': out.push_back('"'); break;
      case '\\': out.push_back('\\'); break;
      case '/': out.push_back('/'); break;
      case 'b': out.push_back('\b'); break;
      case 'f': out.push_back('\f'); break;
      case 'n': out.push_back('\n'); break;
      case 'r': out.push_back('\r'); break;
      case 't': out.push_back('\t'); break;
      case 'u':
      {
         if ( i + 4 > s.size() ) return false;
         uint32_t cp = (uint32_t)strtoul(s.substr(i, 4).c_str(), nullptr, 16);
         i += 4;
         if ( cp >= 0xd800 && cp < 0xdc00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u' )
         {
            uint32_t lo = (uint32_t)strtoul(s.substr(i + 2, 4).c_str(), nullptr, 16);
            i += 6;
            cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
         }
         put_utf8(out, cp);
         break;
      }
      default: return false;
      }
   }
   return false;
}

void json_skip_ws(const std::string& s, size_t& i)
{
   while ( i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r') ) ++i;
}

/* Skip one JSON value (string, number, literal, array or object). */
bool json_skip_value(const std::string& s, size_t& i)
{
   json_skip_ws(s, i);
   if ( i >= s.size() ) return false;
   std::string tmp;
   if ( s[i] == '"' ) return json_read_string(s, i, tmp);
   if ( s[i] == '[' || s[i] == '{' )
   {
      int depth = 0;
      while ( i < s.size() )
      {
         char c = s[i];
         if ( c == '"' )
         {
            if ( !json_read_string(s, i, tmp) ) return false;
            continue;
         }
         if ( c == '[' || c == '{' ) ++depth;
         if ( c == ']' || c == '}' ) --depth;
         ++i;
         if ( depth == 0 ) return true;
      }
      return false;
   }
   while ( i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']' ) ++i;
   return true;
}

/* Value of a top-level key of a JSON object: an array of strings (found =
 * true) or a string (as a one-element vector). */
bool json_top_value(const std::string& s, const std::string& key, std::vector<std::string>& out)
{
   out.clear();
   size_t i = 0;
   json_skip_ws(s, i);
   if ( i >= s.size() || s[i] != '{' ) return false;
   ++i;
   std::string k;
   for ( ;; )
   {
      json_skip_ws(s, i);
      if ( i >= s.size() || s[i] == '}' ) return false;
      if ( !json_read_string(s, i, k) ) return false;
      json_skip_ws(s, i);
      if ( i >= s.size() || s[i] != ':' ) return false;
      ++i;
      json_skip_ws(s, i);
      if ( k == key )
      {
         std::string v;
         if ( i < s.size() && s[i] == '"' )
         {
            if ( !json_read_string(s, i, v) ) return false;
            out.push_back(v);
            return true;
         }
         if ( i >= s.size() || s[i] != '[' ) return false;
         ++i;
         for ( ;; )
         {
            json_skip_ws(s, i);
            if ( i < s.size() && s[i] == ']' ) return true;
            if ( i < s.size() && s[i] == '"' )
            {
               if ( !json_read_string(s, i, v) ) return false;
               out.push_back(v);
            }
            else
            {
               /* a non-string element: keep its text */
               size_t j = i;
               if ( !json_skip_value(s, i) ) return false;
               out.push_back(s.substr(j, i - j));
            }
            json_skip_ws(s, i);
            if ( i < s.size() && s[i] == ',' ) ++i;
         }
      }
      if ( !json_skip_value(s, i) ) return false;
      json_skip_ws(s, i);
      if ( i < s.size() && s[i] == ',' ) ++i;
   }
}

int read_whole_file(const char* path, std::string& buf)
{
   FILE* f = fopen(path, "rb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open");
   std::vector<char> chunk(1 << 20);
   buf.clear();
   size_t n;
   while ( (n = fread(chunk.data(), 1, chunk.size(), f)) > 0 ) buf.append(chunk.data(), n);
   const bool bad = ferror(f) != 0;
   fclose(f);
   if ( bad ) return ferr(KINC_E_INVALID, std::string(path) + ": read failed");
   return KINC_OK;
}

struct ParseError
{
   int64_t line = -1;
   std::string msg;
};

} // namespace

extern "C" int kinc_emx_import_text(const char* text_path, const char* emx_path, const char* nan_token,
                                    int32_t sample_size, int32_t transform, int32_t threads,
                                    int32_t* gene_size_out, int32_t* sample_size_out)
{
   if ( !text_path || !emx_path || sample_size < 0 || transform < KINC_TRANSFORM_NONE
        || transform > KINC_TRANSFORM_LOG10 )
   {
      return ferr(KINC_E_INVALID, "Did not get valid input and/or output arguments.");
   }
   std::string text;
   int rc = read_whole_file(text_path, text);
   if ( rc ) return rc;
   const std::string token = nan_token ? nan_token : "";

   /* QTextStream::readLine: lines end at '\n'; a final line without one
    * still counts, an empty remainder after the last '\n' does not */
   std::vector<std::pair<size_t, size_t>> lines;
   {
      size_t p = 0;
      const size_t n = text.size();
      while ( p < n )
      {
         const void* nl = memchr(text.data() + p, '\n', n - p);
         size_t q = nl ? (size_t)((const char*)nl - text.data()) : n;
         lines.push_back({p, q});
         p = q + 1;
      }
   }

   int S = sample_size;
   std::vector<std::string> samples;
   size_t first = 0;
   std::vector<Word> words;
   if ( S != 0 )
   {
      for ( int i = 0; i < S; ++i ) samples.push_back(std::to_string(i));
   }
   else
   {
      /* the first line(s) until one has words: the sample names */
      while ( first < lines.size() && S == 0 )
      {
         split_words(text.data() + lines[first].first, text.data() + lines[first].second, words);
         S = (int)words.size();
         for ( const Word& w : words ) samples.push_back(w.str());
         ++first;
      }
   }
   const int64_t G64 = (int64_t)(lines.size() - first);
   if ( S < 1 || G64 < 1 ) return ferr(KINC_E_INVALID, std::string(text_path) + ": no expression data");
   if ( G64 > INT32_MAX ) return ferr(KINC_E_INVALID, std::string(text_path) + ": too many genes");
   const int G = (int)G64;

   std::vector<float> data((size_t)G * (size_t)S);
   std::vector<std::string> genes((size_t)G);

   int T = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
   T = std::max(1, std::min(T, 16));
   T = std::min(T, std::max(1, G / 256));
   std::vector<ParseError> errs((size_t)T);

   auto work = [&](int t) {
      const int g0 = (int)((int64_t)G * t / T), g1 = (int)((int64_t)G * (t + 1) / T);
      std::vector<Word> w;
      for ( int g = g0; g < g1; ++g )
      {
         const auto& ln = lines[first + (size_t)g];
         split_words(text.data() + ln.first, text.data() + ln.second, w);
         if ( (int)w.size() != S + 1 )
         {
            errs[t].line = (int64_t)g;
            errs[t].msg = "Encountered gene expression line with incorrect amount of fields. Read in "
                          + std::to_string((long long)w.size() - 1) + " fields when it should have been "
                          + std::to_string(S) + ". Gene name is " + (w.empty() ? std::string() : w[0].str())
                          + ".";
            return;
         }
         genes[(size_t)g] = w[0].str();
         float* row = data.data() + (size_t)g * (size_t)S;
         for ( int i = 1; i <= S; ++i )
         {
            if ( !token.empty() && w[i].equals(token) )
            {
               row[i - 1] = NAN;
               continue;
            }
            double v;
            if ( !parse_double(w[i].b, w[i].e, v) )
            {
               errs[t].line = (int64_t)g;
               errs[t].msg = "Failed to read expression value \"" + w[i].str() + "\" for gene "
                             + w[0].str() + ".";
               return;
            }
            switch ( transform )
            {
            case KINC_TRANSFORM_NLOG: v = std::log(v); break;
            case KINC_TRANSFORM_LOG2: v = std::log2(v); break;
            case KINC_TRANSFORM_LOG10: v = std::log10(v); break;
            default: break;
            }
            row[i - 1] = (float)v;
         }
      }
   };
   if ( T == 1 ) work(0);
   else
   {
      std::vector<std::thread> pool;
      for ( int t = 0; t < T; ++t ) pool.emplace_back(work, t);
      for ( auto& th : pool ) th.join();
   }
   /* the reference stops at the first bad line in file order */
   for ( const ParseError& e : errs )
   {
      if ( e.line >= 0 ) return ferr(KINC_E_INVALID, e.msg);
   }

   std::string meta = "{\"genes\": ";
   json_array(meta, genes);
   meta += ", \"samples\": ";
   json_array(meta, samples);
   meta += ", \"transform\": ";
   json_string(meta, kTransformNames[transform]);
   meta += "}";
   rc = kinc_emx_write(emx_path, data.data(), G, S, meta.c_str());
   if ( rc ) return rc;
   if ( gene_size_out ) *gene_size_out = G;
   if ( sample_size_out ) *sample_size_out = S;
   return KINC_OK;
}

extern "C" int kinc_emx_export_text(const char* emx_path, const char* text_path, const char* nan_token)
{
   if ( !emx_path || !text_path ) return ferr(KINC_E_INVALID, "Did not get valid input and/or output arguments.");
   int32_t G = 0, S = 0;
   int rc = kinc_emx_info(emx_path, &G, &S);
   if ( rc ) return rc;
   std::vector<float> data((size_t)G * (size_t)S);
   rc = kinc_emx_read(emx_path, data.data());
   if ( rc ) return rc;
   int64_t msize = kinc_file_metadata(emx_path, nullptr, 0);
   if ( msize < 0 ) return (int)msize;
   std::string meta((size_t)msize, '\0');
   if ( msize > 0 && kinc_file_metadata(emx_path, &meta[0], msize) < 0 ) return KINC_E_INVALID;

   std::vector<std::string> genes, samples, tname;
   if ( !json_top_value(meta, "genes", genes) || (int)genes.size() != G )
   {
      genes.clear();
      for ( int i = 0; i < G; ++i ) genes.push_back(std::to_string(i));
   }
   if ( !json_top_value(meta, "samples", samples) || (int)samples.size() != S )
   {
      samples.clear();
      for ( int i = 0; i < S; ++i ) samples.push_back(std::to_string(i));
   }
   int transform = KINC_TRANSFORM_NONE;
   if ( json_top_value(meta, "transform", tname) && tname.size() == 1 )
   {
      for ( int t = 0; t < 4; ++t )
      {
         if ( tname[0] == kTransformNames[t] ) transform = t;
      }
   }
   const std::string token = nan_token ? nan_token : "";

   /* rows are formatted by several threads into per-thread buffers of
    * about 8 MB, which are written in gene order */
   int T = (int)std::thread::hardware_concurrency();
   T = std::max(1, std::min(T, 16));
   const int rows_per_chunk = std::max(1, (int)((8u << 20) / ((size_t)S * 14 + 16)));
   auto format_rows = [&](int g0, int g1, std::string& out) {
      char num[64];
      out.clear();
      for ( int g = g0; g < g1; ++g )
      {
         out += genes[(size_t)g];
         const float* row = data.data() + (size_t)g * (size_t)S;
         for ( int j = 0; j < S; ++j )
         {
            float value = row[j];
            out.push_back('\t');
            if ( std::isnan(value) )
            {
               out += token;
               continue;
            }
            switch ( transform )
            {
            case KINC_TRANSFORM_NLOG: value = std::exp(value); break;
            case KINC_TRANSFORM_LOG2: value = (float)std::pow(2.0, (double)value); break;
            case KINC_TRANSFORM_LOG10: value = (float)std::pow(10.0, (double)value); break;
            default: break;
            }
            snprintf(num, sizeof(num), "%.12g", (double)value);
            out += num;
         }
         out.push_back('\n');
      }
   };

   FILE* f = fopen(text_path, "wb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(text_path) + ": cannot open for writing");
   std::string head;
   for ( int i = 0; i < S; ++i )
   {
      head += samples[(size_t)i];
      head.push_back('\t');
   }
   head.push_back('\n');
   bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
   std::vector<std::string> bufs((size_t)T);
   for ( int g = 0; g < G && ok; g += T * rows_per_chunk )
   {
      std::vector<std::thread> pool;
      int used = 0;
      for ( int t = 0; t < T; ++t )
      {
         const int g0 = g + t * rows_per_chunk;
         if ( g0 >= G ) break;
         const int g1 = std::min(G, g0 + rows_per_chunk);
         pool.emplace_back(format_rows, g0, g1, std::ref(bufs[(size_t)t]));
         ++used;
      }
      for ( auto& th : pool ) th.join();
      for ( int t = 0; t < used && ok; ++t )
      {
         ok = fwrite(bufs[(size_t)t].data(), 1, bufs[(size_t)t].size(), f) == bufs[(size_t)t].size();
      }
   }
   ok = (fclose(f) == 0) && ok;
   return ok ? KINC_OK : ferr(KINC_E_INVALID, std::string(text_path) + ": write failed");
}

/* ------------------------------------------- CMX text export and extract */

namespace {

/* A pairwise file (CCM or CMX) read whole: header fields and records. */
struct PairwiseIn
{
   int32_t G = 0, max_clusters = 0, data_size = 0;
   int64_t pairs = 0, clusters = 0;
   std::vector<uint8_t> sub;
   std::vector<uint8_t> payload;
   std::string meta;
   size_t base = 0;   /* offset of the first record in payload */

   int32_t x(int64_t i) const { return (int32_t)get32(rec(i)); }
   int32_t y(int64_t i) const { return (int32_t)get32(rec(i) + 4); }
   int cluster(int64_t i) const { return (int8_t)rec(i)[8]; }
   const uint8_t* data(int64_t i) const { return rec(i) + 9; }
   const uint8_t* rec(int64_t i) const { return payload.data() + base + (size_t)i * (9 + (size_t)data_size); }
   int64_t pair_index(int64_t i) const { return (int64_t)x(i) * (x(i) - 1) / 2 + y(i); }
};

int read_pairwise(const char* path, uint32_t type, PairwiseIn& p)
{
   FILE* f = fopen(path, "rb");
   if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open");
   EnvelopeInfo e;
   int rc = read_envelope(f, e, path);
   if ( rc == KINC_OK && e.type != type )
      rc = ferr(KINC_E_INVALID, std::string(path) + (type == KINC_FILE_CCM ? ": not a CCM file" : ": not a CMX file"));
   if ( rc == KINC_OK )
   {
      p.payload.resize((size_t)e.payload_size);
      p.meta.resize((size_t)e.meta_size);
      if ( fseek(f, (long)e.payload_off, SEEK_SET) != 0
           || fread(p.payload.data(), 1, p.payload.size(), f) != p.payload.size()
           || fseek(f, (long)e.meta_off, SEEK_SET) != 0
           || (p.meta.size() && fread(&p.meta[0], 1, p.meta.size(), f) != p.meta.size()) )
      {
         rc = ferr(KINC_E_INVALID, std::string(path) + ": short file");
      }
   }
   fclose(f);
   if ( rc ) return rc;
   const uint8_t* h = p.payload.data();
   if ( p.payload.size() < 30 ) return ferr(KINC_E_INVALID, std::string(path) + ": short header");
   p.G = (int32_t)get32(h);
   p.max_clusters = (int32_t)get32(h + 4);
   p.data_size = (int32_t)get32(h + 8);
   p.pairs = (int64_t)get64(h + 12);
   p.clusters = (int64_t)get64(h + 20);
   const size_t off = (size_t)h[28] | ((size_t)h[29] << 8);
   p.base = 30 + off;
   if ( p.base > p.payload.size() || p.data_size < 0
        || p.base + (size_t)p.clusters * (9 + (size_t)p.data_size) > p.payload.size() )
   {
      return ferr(KINC_E_INVALID, std::string(path) + ": truncated records");
   }
   p.sub.assign(h + 30, h + p.base);
   return KINC_OK;
}

/* CCM mask nibble of sample i (CCMatrix::Pair, ccmatrix.cpp:223-243) */
inline int ccm_nibble(const uint8_t* data, int i)
{
   const uint8_t b = data[i / 2];
   return (i & 1) ? (b >> 4) : (b & 0x0f);
}

struct MaskStats
{
   int samples = 0, missing = 0, post = 0, pre = 0, threshold = 0;
};

/* the summary counts and the mask text of one CCM cluster */
void ccm_cluster(const uint8_t* data, int S, MaskStats& st, std::string& mask)
{
   st = MaskStats();
   mask.resize((size_t)S);
   for ( int i = 0; i < S; ++i )
   {
      const int v = ccm_nibble(data, i);
      switch ( v )
      {
      case 1: ++st.samples; break;
      case 6: ++st.threshold; break;
      case 7: ++st.pre; break;
      case 8: ++st.post; break;
      case 9: ++st.missing; break;
      default: break;
      }
      mask[(size_t)i] = (char)('0' + v);
   }
}

/* Walks the CMX pairs in file order with the CCM record of each of their
 * clusters (both files hold the pairs in increasing index order). */
struct PairWalker
{
   const PairwiseIn& cmx;
   const PairwiseIn& ccm;
   int64_t i = 0;      /* first CMX record of the current pair */
   int64_t n = 0;      /* its cluster count */
   int64_t j = 0;      /* CCM cursor */
   int64_t ccm_first = -1;

   PairWalker(const PairwiseIn& a, const PairwiseIn& b) : cmx(a), ccm(b) {}

   bool next()
   {
      i += n;
      if ( i >= cmx.clusters ) return false;
      const int64_t idx = cmx.pair_index(i);
      n = 1;
      while ( i + n < cmx.clusters && cmx.pair_index(i + n) == idx ) ++n;
      ccm_first = -1;
      if ( n > 1 )
      {
         while ( j < ccm.clusters && ccm.pair_index(j) < idx ) ++j;
         if ( j < ccm.clusters && ccm.pair_index(j) == idx ) ccm_first = j;
      }
      return true;
   }

   /* the CCM data of cluster k of the current pair, or null */
   const uint8_t* ccm_data(int k) const
   {
      if ( ccm_first < 0 || ccm_first + k >= ccm.clusters ) return nullptr;
      if ( ccm.pair_index(ccm_first + k) != cmx.pair_index(i) || ccm.cluster(ccm_first + k) != k ) return nullptr;
      return ccm.data(ccm_first + k);
   }
};

float cmx_corr(const PairwiseIn& cmx, int64_t r)
{
   float c;
   memcpy(&c, cmx.data(r), 4);
   return c;
}

/* buffered text output */
struct TextOut
{
   FILE* f = nullptr;
   std::string buf;
   bool ok = true;
   int open(const char* path)
   {
      f = fopen(path, "wb");
      if ( !f ) return ferr(KINC_E_INVALID, std::string(path) + ": cannot open for writing");
      buf.reserve(1 << 20);
      return KINC_OK;
   }
   void put(const std::string& s)
   {
      buf += s;
      if ( buf.size() >= (1u << 20) ) flush();
   }
   void flush()
   {
      if ( ok && !buf.empty() ) ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
      buf.clear();
   }
   bool close()
   {
      flush();
      const bool c = f && fclose(f) == 0;
      f = nullptr;
      return ok && c;
   }
};

std::string fmt_g(float v)
{
   char b[32];
   snprintf(b, sizeof(b), "%g", (double)v);
   return b;
}

int ccm_samples(const PairwiseIn& ccm)
{
   return ccm.sub.size() >= 4 ? (int)get32(ccm.sub.data()) : 0;
}

} // namespace

extern "C" int kinc_export_cmx_text(const char* ccm_path, const char* cmx_path, const char* out_path)
{
   if ( !ccm_path || !cmx_path || !out_path )
      return ferr(KINC_E_INVALID, "Did not get valid input and/or output arguments.");
   PairwiseIn ccm, cmx;
   int rc = read_pairwise(ccm_path, KINC_FILE_CCM, ccm);
   if ( rc == KINC_OK ) rc = read_pairwise(cmx_path, KINC_FILE_CMX, cmx);
   if ( rc ) return rc;
   const int S = ccm_samples(ccm);
   TextOut out;
   rc = out.open(out_path);
   if ( rc ) return rc;
   PairWalker w(cmx, ccm);
   std::string mask, line;
   MaskStats st;
   while ( w.next() )
   {
      for ( int k = 0; k < (int)w.n; ++k )
      {
         const float corr = cmx_corr(cmx, w.i + k);
         st = MaskStats();
         if ( w.n > 1 )
         {
            const uint8_t* d = w.ccm_data(k);
            if ( !d )
            {
               out.close();
               return ferr(KINC_E_INVALID, "CCM has no record for pair (" + std::to_string(cmx.x(w.i)) + ", "
                                              + std::to_string(cmx.y(w.i)) + ") cluster " + std::to_string(k));
            }
            ccm_cluster(d, S, st, mask);
         }
         else
         {
            mask.assign((size_t)S, '0');
         }
         line = std::to_string(cmx.x(w.i)) + "\t" + std::to_string(cmx.y(w.i)) + "\t" + std::to_string(k) + "\t"
                + std::to_string(w.n) + "\t" + std::to_string(st.samples) + "\t" + std::to_string(st.missing)
                + "\t" + std::to_string(st.post) + "\t" + std::to_string(st.pre) + "\t"
                + std::to_string(st.threshold) + "\t" + fmt_g(corr) + "\t" + mask + "\n";
         out.put(line);
      }
   }
   return out.close() ? KINC_OK : ferr(KINC_E_INVALID, std::string(out_path) + ": write failed");
}

extern "C" int kinc_extract_network(const char* emx_path, const char* ccm_path, const char* cmx_path,
                                    const char* out_path, const char* graphml_path, float min_corr,
                                    float max_corr)
{
   if ( !emx_path || !ccm_path || !cmx_path || !out_path )
      return ferr(KINC_E_INVALID, "Did not get valid input and/or output arguments.");
   PairwiseIn ccm, cmx;
   int rc = read_pairwise(ccm_path, KINC_FILE_CCM, ccm);
   if ( rc == KINC_OK ) rc = read_pairwise(cmx_path, KINC_FILE_CMX, cmx);
   if ( rc ) return rc;
   int32_t G = 0, SE = 0;
   rc = kinc_emx_info(emx_path, &G, &SE);
   if ( rc ) return rc;
   std::vector<float> emx((size_t)G * (size_t)SE);
   rc = kinc_emx_read(emx_path, emx.data());
   if ( rc ) return rc;
   const int S = ccm_samples(ccm);
   if ( SE != S ) return ferr(KINC_E_INVALID, "the expression matrix and the CCM have different sample counts");
   std::vector<std::string> names;
   if ( !json_top_value(cmx.meta, "genes", names) || (int)names.size() < cmx.G )
   {
      names.clear();
      for ( int g = 0; g < cmx.G; ++g ) names.push_back(std::to_string(g));
   }
   if ( cmx.G > G ) return ferr(KINC_E_INVALID, "the CMX has more genes than the expression matrix");

   /* the mask of one cluster: CCM data for a multi-cluster pair, else the
    * NaN pattern of the two genes */
   auto cluster_mask = [&](const PairWalker& w, int k, MaskStats& st, std::string& mask) -> bool {
      st = MaskStats();
      if ( w.n > 1 )
      {
         const uint8_t* d = w.ccm_data(k);
         if ( !d ) return false;
         ccm_cluster(d, S, st, mask);
         return true;
      }
      const float* a = emx.data() + (size_t)cmx.x(w.i) * S;
      const float* b = emx.data() + (size_t)cmx.y(w.i) * S;
      mask.resize((size_t)S);
      for ( int i = 0; i < S; ++i ) mask[(size_t)i] = (std::isnan(a[i]) || std::isnan(b[i])) ? '9' : '1';
      return true;
   };
   auto excluded = [&](float c) { return std::fabs(c) < min_corr || max_corr < std::fabs(c); };

   TextOut out;
   rc = out.open(out_path);
   if ( rc ) return rc;
   out.put("Source\tTarget\tsc\tInteraction\tCluster\tNum_Clusters\tCluster_Samples\tMissing_Samples\t"
           "Cluster_Outliers\tPair_Outliers\tToo_Low\tSamples\n");
   std::string mask, line;
   MaskStats st;
   {
      PairWalker w(cmx, ccm);
      while ( w.next() )
      {
         for ( int k = 0; k < (int)w.n; ++k )
         {
            const float corr = cmx_corr(cmx, w.i + k);
            if ( excluded(corr) ) continue;
            if ( !cluster_mask(w, k, st, mask) )
            {
               out.close();
               return ferr(KINC_E_INVALID, "CCM has no record for a multi-cluster pair");
            }
            line = names[(size_t)cmx.x(w.i)] + "\t" + names[(size_t)cmx.y(w.i)] + "\t" + fmt_g(corr) + "\tco\t"
                   + std::to_string(k) + "\t" + std::to_string(w.n) + "\t" + std::to_string(st.samples) + "\t"
                   + std::to_string(st.missing) + "\t" + std::to_string(st.post) + "\t" + std::to_string(st.pre)
                   + "\t" + std::to_string(st.threshold) + "\t" + mask + "\n";
            out.put(line);
         }
      }
   }
   if ( !out.close() ) return ferr(KINC_E_INVALID, std::string(out_path) + ": write failed");
   if ( !graphml_path ) return KINC_OK;

   TextOut gml;
   rc = gml.open(graphml_path);
   if ( rc ) return rc;
   gml.put("<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
           "<graphml xmlns=\"http://graphml.graphdrawing.org/xmlns\"\n"
           "    xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
           "    xsi:schemaLocation=\"http://graphml.graphdrawing.org/xmlns/1.0/graphml.xsd\">\n"
           "  <graph id=\"G\" edgedefault=\"undirected\">\n");
   for ( int g = 0; g < cmx.G; ++g ) gml.put("    <node id=\"" + names[(size_t)g] + "\"/>\n");
   {
      PairWalker w(cmx, ccm);
      while ( w.next() )
      {
         for ( int k = 0; k < (int)w.n; ++k )
         {
            const float corr = cmx_corr(cmx, w.i + k);
            if ( excluded(corr) ) continue;
            if ( !cluster_mask(w, k, st, mask) )
            {
               gml.close();
               return ferr(KINC_E_INVALID, "CCM has no record for a multi-cluster pair");
            }
            gml.put("    <edge source=\"" + names[(size_t)cmx.x(w.i)] + "\" target=\"" + names[(size_t)cmx.y(w.i)]
                    + "\" samples=\"" + mask + "\"/>\n");
         }
      }
   }
   gml.put("  </graph>\n</graphml>\n");
   return gml.close() ? KINC_OK : ferr(KINC_E_INVALID, std::string(graphml_path) + ": write failed");
}
