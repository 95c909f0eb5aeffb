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
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kinc_formats.h"
#include "../../include/kinc_sim.h"

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

struct kinc_similarity_writer
{
   PairwiseFile ccm;
   PairwiseFile cmx;
   int32_t S = 0;
   int32_t max_clusters = 0;
   float min_corr = 0.5f;
   float max_corr = 1.0f;
   std::vector<uint8_t> masks;
   std::vector<float> kept;
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

extern "C" int kinc_similarity_writer_process(kinc_similarity_writer* w, int64_t start, int64_t count,
                                              const int8_t* K, const int8_t* labels,
                                              const float* corr, int32_t corr_stride)
{
   if ( !w || !K || !corr || count < 0 ) return ferr(KINC_E_INVALID, "invalid process arguments");
   const int S = w->S;
   const int nb = (S + 1) / 2;
   int32_t x, y;
   kinc_pair_index(start, &x, &y);
   for ( int64_t q = 0; q < count; ++q )
   {
      const int k_sel = K[q];
      const float* r = corr + q * corr_stride;
      if ( k_sel > w->max_clusters )
      {
         return ferr(KINC_E_INVALID, "Cannot write pair with cluster size exceeding the max");
      }
      if ( k_sel > 1 )
      {
         if ( !labels ) return ferr(KINC_E_INVALID, "labels are required when K > 1");
         const int8_t* lab = labels + q * S;
         w->masks.clear();
         int n = 0;
         for ( int k = 0; k < k_sel; ++k )
         {
            const float c = r[k];
            if ( !std::isnan(c) && w->min_corr <= std::fabs(c) && std::fabs(c) <= w->max_corr )
            {
               size_t base = w->masks.size();
               w->masks.resize(base + (size_t)nb, 0);
               for ( int i = 0; i < S; ++i )
               {
                  const int v = (lab[i] >= 0) ? (k == lab[i]) : -lab[i];
                  const uint8_t nib = (uint8_t)(v & 0x0f);
                  w->masks[base + (size_t)(i / 2)] |= (i & 1) ? (uint8_t)(nib << 4) : nib;
               }
               ++n;
            }
         }
         for ( int c = 0; c < n; ++c )
         {
            int rc = w->ccm.write(x, y, c, w->masks.data() + (size_t)c * (size_t)nb);
            if ( rc ) return rc;
         }
         if ( n > 0 ) ++w->ccm.pair_size;
      }
      if ( k_sel > 0 )
      {
         int n = 0;
         for ( int k = 0; k < k_sel; ++k )
         {
            const float c = r[k];
            if ( !std::isnan(c) && w->min_corr <= std::fabs(c) && std::fabs(c) <= w->max_corr )
            {
               uint8_t b[4];
               memcpy(b, &c, 4);
               int rc = w->cmx.write(x, y, n, b);
               if ( rc ) return rc;
               ++n;
            }
         }
         if ( n > 0 ) ++w->cmx.pair_size;
      }
      /* Pairwise::Index::operator++ */
      if ( ++y >= x )
      {
         y = 0;
         ++x;
      }
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
