/*
 * kinc_cli.cpp — `kinc run similarity`, the command-line entry of the
 * Similarity analytic on one MI355X.
 *
 * Mirrors the reference's command (src/core/similarity_input.cpp:96-254,
 * defaults src/core/similarity.h:55-71):
 *   kinc run similarity --input A.emx --ccm A.ccm --cmx A.cmx
 *        [--clusmethod none|gmm|kmeans] [--corrmethod pearson|spearman]
 *        [--minexpr F] [--minsamp N] [--minclus N] [--maxclus N]
 *        [--crit BIC|ICL] [--preout true|false] [--postout true|false]
 *        [--mincorr F] [--maxcorr F] [--ksize N]
 * plus flags the reference does not have: --device N, --form
 * plain|persistent, --pearson-fp64 BOOL (Pearson sums and quotient in
 * double), --workers 1|2 (engine handles taking blocks in turn) and
 * --autotune BOOL (default true: time the launch configurations on the first
 * work block with all worker handles, kinc_sim_autotune_group; results are
 * the same either way); and the
 * expression-matrix converters around it (importexpressionmatrix_input.cpp,
 * exportexpressionmatrix_input.cpp):
 *   kinc run import-emx --input A.txt --output A.emx [--nan TOKEN]
 *        [--samples N] [--transform none|"natural logarithm"|...] [--threads N]
 *   kinc run export-emx --input A.emx --output A.txt [--nan TOKEN]
 * the correlation-matrix exporters (exportcorrelationmatrix_input.cpp,
 * extract_input.cpp):
 *   kinc run export-cmx --ccm A.ccm --cmx A.cmx --output A.txt
 *   kinc run extract --emx A.emx --ccm A.ccm --cmx A.cmx --output net.txt
 *        [--graphml net.graphml] [--mincorr 0.85] [--maxcorr 1]
 * and the RMT threshold analytic (rmt_input.cpp:58-136):
 *   kinc run rmt --input A.cmx --log A.log [--tstart F] [--tstep F]
 *        [--tstop F] [--minpace N] [--maxpace N] [--bins N] [--device N]
 *
 * The ACE driver loop is Similarity::size/makeWork (similarity.cpp:17-39)
 * then execute and process in pair order.  Here the device executes block
 * b+1 while the host writes block b: two device output sets and two pinned
 * host sets alternate, and process() still sees the blocks in order.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kinc_formats.h"
#include "../../include/kinc_rmt.h"
#include "../../include/kinc_sim.h"

namespace {

bool parse_bool(const std::string& v)
{
   return v == "true" || v == "1" || v == "yes" || v == "on";
}

int usage()
{
   fprintf(stderr,
           "usage: kinc run similarity --input EMX --ccm CCM --cmx CMX [--clusmethod none|gmm|kmeans]\n"
           "       [--corrmethod pearson|spearman] [--minexpr F] [--minsamp N] [--minclus N]\n"
           "       [--maxclus N] [--crit BIC|ICL] [--preout BOOL] [--postout BOOL]\n"
           "       [--mincorr F] [--maxcorr F] [--ksize N] [--device N] [--form plain|persistent]\n"
           "       [--pearson-fp64 BOOL] [--workers 1|2] [--autotune BOOL]\n"
           "       kinc run import-emx --input TXT --output EMX [--nan TOKEN] [--samples N]\n"
           "       [--transform none|'natural logarithm'|'logarithm base 2'|'logarithm base 10']\n"
           "       [--threads N]\n"
           "       kinc run export-emx --input EMX --output TXT [--nan TOKEN]\n"
           "       kinc run export-cmx --ccm CCM --cmx CMX --output TXT\n"
           "       kinc run extract --emx EMX --ccm CCM --cmx CMX --output TXT [--graphml GML]\n"
           "       [--mincorr F] [--maxcorr F]\n"
           "       kinc run rmt --input CMX --log LOG [--tstart F] [--tstep F] [--tstop F]\n"
           "       [--minpace N] [--maxpace N] [--bins N] [--device N]\n");
   return 2;
}

/* export-cmx / extract */
int run_export(bool extract, int argc, char** argv)
{
   std::string emx, ccm, cmx, out, gml;
   float mincorr = 0.85f, maxcorr = 1.0f;
   for ( int i = 3; i < argc; i += 2 )
   {
      if ( i + 1 >= argc ) return usage();
      const std::string k = argv[i], v = argv[i + 1];
      if ( k == "--ccm" ) ccm = v;
      else if ( k == "--cmx" ) cmx = v;
      else if ( k == "--output" ) out = v;
      else if ( extract && k == "--emx" ) emx = v;
      else if ( extract && k == "--graphml" ) gml = v;
      else if ( extract && k == "--mincorr" ) mincorr = strtof(v.c_str(), nullptr);
      else if ( extract && k == "--maxcorr" ) maxcorr = strtof(v.c_str(), nullptr);
      else return usage();
   }
   if ( ccm.empty() || cmx.empty() || out.empty() || (extract && emx.empty()) )
   {
      fprintf(stderr, "Did not get valid input and/or output arguments.\n");
      return 2;
   }
   const int rc = extract ? kinc_extract_network(emx.c_str(), ccm.c_str(), cmx.c_str(), out.c_str(),
                                                 gml.empty() ? nullptr : gml.c_str(), mincorr, maxcorr)
                          : kinc_export_cmx_text(ccm.c_str(), cmx.c_str(), out.c_str());
   if ( rc != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      return 1;
   }
   return 0;
}

int run_rmt(int argc, char** argv)
{
   kinc_rmt_params p;
   kinc_rmt_default_params(&p);
   std::string in, log;
   int device = 0;
   for ( int i = 3; i < argc; i += 2 )
   {
      if ( i + 1 >= argc ) return usage();
      const std::string k = argv[i], v = argv[i + 1];
      if ( k == "--input" ) in = v;
      else if ( k == "--log" ) log = v;
      else if ( k == "--tstart" ) p.threshold_start = strtof(v.c_str(), nullptr);
      else if ( k == "--tstep" ) p.threshold_step = strtof(v.c_str(), nullptr);
      else if ( k == "--tstop" ) p.threshold_stop = strtof(v.c_str(), nullptr);
      else if ( k == "--minpace" ) p.min_pace = atoi(v.c_str());
      else if ( k == "--maxpace" ) p.max_pace = atoi(v.c_str());
      else if ( k == "--bins" ) p.bins = atoi(v.c_str());
      else if ( k == "--device" ) device = atoi(v.c_str());
      else return usage();
   }
   if ( in.empty() || log.empty() )
   {
      fprintf(stderr, "Did not get valid input or logfile arguments.\n");
      return 2;
   }
   float t = 0;
   auto t0 = std::chrono::steady_clock::now();
   if ( kinc_rmt_run_file(&p, in.c_str(), log.c_str(), device, &t) != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_rmt_last_error(nullptr));
      return 1;
   }
   double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   fprintf(stderr, "threshold %g in %.2f s\n", (double)t, s);
   return 0;
}

/* import-emx / export-emx */
int run_convert(bool import, int argc, char** argv)
{
   static const char* kNames[4] = {"none", "natural logarithm", "logarithm base 2", "logarithm base 10"};
   std::string in, out, nan;
   int samples = 0, transform = KINC_TRANSFORM_NONE, threads = 0;
   for ( int i = 3; i < argc; i += 2 )
   {
      if ( i + 1 >= argc ) return usage();
      const std::string k = argv[i], v = argv[i + 1];
      if ( k == "--input" ) in = v;
      else if ( k == "--output" ) out = v;
      else if ( k == "--nan" ) nan = v;
      else if ( import && k == "--samples" ) samples = atoi(v.c_str());
      else if ( import && k == "--threads" ) threads = atoi(v.c_str());
      else if ( import && k == "--transform" )
      {
         transform = -1;
         for ( int t = 0; t < 4; ++t )
         {
            if ( v == kNames[t] ) transform = t;
         }
         if ( transform < 0 ) return usage();
      }
      else return usage();
   }
   if ( in.empty() || out.empty() )
   {
      fprintf(stderr, "Did not get valid input and/or output arguments.\n");
      return 2;
   }
   int rc;
   if ( import )
   {
      int32_t G = 0, S = 0;
      auto t0 = std::chrono::steady_clock::now();
      rc = kinc_emx_import_text(in.c_str(), out.c_str(), nan.c_str(), samples, transform, threads, &G, &S);
      if ( rc == KINC_OK )
      {
         double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
         fprintf(stderr, "imported %d genes x %d samples in %.2f s\n", G, S, s);
      }
   }
   else
   {
      rc = kinc_emx_export_text(in.c_str(), out.c_str(), nan.c_str());
   }
   if ( rc != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      return 1;
   }
   return 0;
}

struct HostSet
{
   int8_t* K = nullptr;
   int8_t* labels = nullptr;
   float* corr = nullptr;
};

struct DevSet
{
   int8_t* K = nullptr;
   int8_t* labels = nullptr;
   float* corr = nullptr;
};

} // namespace

int main(int argc, char** argv)
{
   if ( argc < 3 || std::string(argv[1]) != "run" ) return usage();
   if ( std::string(argv[2]) == "import-emx" ) return run_convert(true, argc, argv);
   if ( std::string(argv[2]) == "export-emx" ) return run_convert(false, argc, argv);
   if ( std::string(argv[2]) == "rmt" ) return run_rmt(argc, argv);
   if ( std::string(argv[2]) == "export-cmx" ) return run_export(false, argc, argv);
   if ( std::string(argv[2]) == "extract" ) return run_export(true, argc, argv);
   if ( std::string(argv[2]) != "similarity" ) return usage();
   kinc_sim_params p;
   kinc_sim_default_params(&p);
   std::string in, ccm, cmx;
   int device = 0;
   int workers = 2;
   int autotune = 1;
   for ( int i = 3; i < argc; i += 2 )
   {
      if ( i + 1 >= argc ) return usage();
      const std::string k = argv[i], v = argv[i + 1];
      if ( k == "--input" ) in = v;
      else if ( k == "--ccm" ) ccm = v;
      else if ( k == "--cmx" ) cmx = v;
      else if ( k == "--clusmethod" ) p.clus_method = v == "gmm" ? KINC_CLUS_GMM : v == "kmeans" ? KINC_CLUS_KMEANS : KINC_CLUS_NONE;
      else if ( k == "--corrmethod" ) p.corr_method = v == "spearman" ? KINC_CORR_SPEARMAN : KINC_CORR_PEARSON;
      else if ( k == "--minexpr" ) p.min_expr = strtof(v.c_str(), nullptr);
      else if ( k == "--minsamp" ) p.min_samples = atoi(v.c_str());
      else if ( k == "--minclus" ) p.min_clusters = atoi(v.c_str());
      else if ( k == "--maxclus" ) p.max_clusters = atoi(v.c_str());
      else if ( k == "--crit" ) p.criterion = v == "BIC" ? KINC_CRIT_BIC : KINC_CRIT_ICL;
      else if ( k == "--preout" ) p.pre_out = parse_bool(v);
      else if ( k == "--postout" ) p.post_out = parse_bool(v);
      else if ( k == "--mincorr" ) p.min_corr = strtof(v.c_str(), nullptr);
      else if ( k == "--maxcorr" ) p.max_corr = strtof(v.c_str(), nullptr);
      else if ( k == "--ksize" ) p.ksize = atoi(v.c_str());
      else if ( k == "--device" ) device = atoi(v.c_str());
      else if ( k == "--workers" ) workers = atoi(v.c_str());
      else if ( k == "--autotune" ) autotune = parse_bool(v);
      else if ( k == "--form" ) p.form = v == "plain" ? KINC_FORM_PLAIN : KINC_FORM_PERSISTENT;
      else if ( k == "--pearson-fp64" ) p.pearson_fp64 = parse_bool(v);
      else return usage();
   }
   if ( in.empty() || ccm.empty() || cmx.empty() )
   {
      fprintf(stderr, "Did not get valid input and/or output arguments.\n");
      return 2;
   }
   if ( workers < 1 || workers > 2 )
   {
      fprintf(stderr, "--workers must be 1 or 2.\n");
      return 2;
   }

   int32_t G = 0, S = 0;
   if ( kinc_emx_info(in.c_str(), &G, &S) != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      return 1;
   }
   std::vector<float> emx((size_t)G * (size_t)S);
   if ( kinc_emx_read(in.c_str(), emx.data()) != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      return 1;
   }
   int64_t msize = kinc_file_metadata(in.c_str(), nullptr, 0);
   std::string meta(msize > 0 ? (size_t)msize : 0, '\0');
   if ( msize > 0 ) kinc_file_metadata(in.c_str(), &meta[0], msize);

   /* blocks go in turn to `workers` engine handles, each with its own
    * stream, so a block's bins start while the previous block's last pairs
    * finish (as the reference runs several workers per device) */
   kinc_sim* hs[2] = {nullptr, nullptr};
   for ( int k = 0; k < workers; ++k )
   {
      if ( kinc_sim_create(&p, emx.data(), G, S, device, &hs[k]) != KINC_OK )
      {
         fprintf(stderr, "%s\n", kinc_sim_last_error(nullptr));
         for ( int j = 0; j < k; ++j ) kinc_sim_destroy(hs[j]);
         return 1;
      }
   }
   /* launch configuration: timed on the first work block with all worker
    * handles running it at once, as they share the device later (results do
    * not depend on it) */
   if ( autotune )
   {
      const int64_t n = std::min<int64_t>(KINC_WORK_BLOCK_SIZE, kinc_pair_count(G));
      kinc_sim_tuning t;
      if ( kinc_sim_autotune_group(hs, workers, 0, n, &t) != KINC_OK )
      {
         fprintf(stderr, "%s\n", kinc_sim_last_error(hs[0]));
         for ( int k = 0; k < workers; ++k ) kinc_sim_destroy(hs[k]);
         return 1;
      }
   }
   kinc_similarity_writer* w = nullptr;
   if ( kinc_similarity_writer_open(ccm.c_str(), cmx.c_str(), G, S, p.max_clusters, p.corr_method,
                                    p.min_corr, p.max_corr, meta.empty() ? nullptr : meta.c_str(), &w) != KINC_OK )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      for ( int k = 0; k < workers; ++k ) kinc_sim_destroy(hs[k]);
      return 1;
   }

   const int64_t total = kinc_pair_count(G);
   const int64_t B = KINC_WORK_BLOCK_SIZE;
   const int64_t nblocks = (total + B - 1) / B;
   const int MC = p.max_clusters;
   hipStream_t streams[2] = {nullptr, nullptr};
   DevSet dev[2];
   HostSet host[2];
   hipEvent_t done[2] = {nullptr, nullptr};
   bool ok = hipSetDevice(device) == hipSuccess;
   for ( int k = 0; ok && k < workers; ++k )
   {
      ok = hipStreamCreateWithFlags(&streams[k], hipStreamNonBlocking) == hipSuccess;
   }
   for ( int s = 0; ok && s < 2; ++s )
   {
      ok = hipMalloc(&dev[s].K, B) == hipSuccess && hipMalloc(&dev[s].labels, B * S) == hipSuccess
        && hipMalloc(&dev[s].corr, sizeof(float) * B * MC) == hipSuccess
        && hipHostMalloc(&host[s].K, B, 0) == hipSuccess && hipHostMalloc(&host[s].labels, B * S, 0) == hipSuccess
        && hipHostMalloc(&host[s].corr, sizeof(float) * B * MC, 0) == hipSuccess
        && hipEventCreate(&done[s]) == hipSuccess;
   }
   int rc = 0;
   if ( !ok )
   {
      fprintf(stderr, "device allocation failed\n");
      rc = 1;
   }

   /* without clustering every pair has K = 1 and the writer needs no labels */
   const bool want_labels = p.clus_method != KINC_CLUS_NONE;
   auto t0 = std::chrono::steady_clock::now();
   auto submit = [&](int64_t b) -> bool {
      const int s = (int)(b & 1);
      kinc_sim* h = hs[s % workers];
      hipStream_t stream = streams[s % workers];
      const int64_t start = b * B;
      const int64_t count = std::min(B, total - start);
      if ( kinc_sim_execute_device(h, start, count, dev[s].K, want_labels ? dev[s].labels : nullptr,
                                   dev[s].corr, nullptr, stream) != KINC_OK )
      {
         fprintf(stderr, "%s\n", kinc_sim_last_error(h));
         return false;
      }
      return hipMemcpyAsync(host[s].K, dev[s].K, count, hipMemcpyDeviceToHost, stream) == hipSuccess
          && (!want_labels
              || hipMemcpyAsync(host[s].labels, dev[s].labels, count * S, hipMemcpyDeviceToHost, stream) == hipSuccess)
          && hipMemcpyAsync(host[s].corr, dev[s].corr, sizeof(float) * count * MC, hipMemcpyDeviceToHost, stream) == hipSuccess
          && hipEventRecord(done[s], stream) == hipSuccess;
   };
   if ( rc == 0 && nblocks > 0 && !submit(0) ) rc = 1;
   for ( int64_t b = 0; rc == 0 && b < nblocks; ++b )
   {
      const int s = (int)(b & 1);
      if ( b + 1 < nblocks && !submit(b + 1) )
      {
         rc = 1;
         break;
      }
      if ( hipEventSynchronize(done[s]) != hipSuccess )
      {
         rc = 1;
         break;
      }
      const int64_t start = b * B;
      const int64_t count = std::min(B, total - start);
      if ( kinc_similarity_writer_process(w, start, count, host[s].K, want_labels ? host[s].labels : nullptr,
                                          host[s].corr, MC) != KINC_OK )
      {
         fprintf(stderr, "%s\n", kinc_formats_last_error());
         rc = 1;
      }
   }
   if ( kinc_similarity_writer_close(w) != KINC_OK && rc == 0 )
   {
      fprintf(stderr, "%s\n", kinc_formats_last_error());
      rc = 1;
   }
   const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   if ( rc == 0 )
   {
      fprintf(stderr, "similarity: %lld pairs in %lld blocks, %.3f s (%.0f pairs/s)\n",
              (long long)total, (long long)nblocks, secs, total / secs);
   }
   /* after an error blocks may still be in flight on the streams */
   for ( int k = 0; k < workers; ++k )
      if ( streams[k] ) (void)hipStreamSynchronize(streams[k]);
   for ( int s = 0; s < 2; ++s )
   {
      (void)hipFree(dev[s].K);
      (void)hipFree(dev[s].labels);
      (void)hipFree(dev[s].corr);
      (void)hipHostFree(host[s].K);
      (void)hipHostFree(host[s].labels);
      (void)hipHostFree(host[s].corr);
      if ( done[s] ) (void)hipEventDestroy(done[s]);
   }
   for ( int k = 0; k < workers; ++k )
   {
      if ( streams[k] ) (void)hipStreamDestroy(streams[k]);
      kinc_sim_destroy(hs[k]);
   }
   return rc;
}
