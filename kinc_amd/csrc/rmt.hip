/*
 * rmt.hip — the RMT threshold analytic on one MI355X (include/kinc_rmt.h).
 *
 * Device side: the correlation matrix as K dense G x G float planes (HBM),
 * and the gather of a pruned block into a double matrix whose eigenvalues
 * rocSOLVER's dsyevd computes.  Host side: the sweep of RMT::process
 * (src/core/rmt.cpp:33-109), the connected components of the pruned matrix,
 * and the chi-square of the eigenvalues (rmt.cpp:282-439).
 *
 * The pruned matrix of the reference (computePruneMatrix, rmt.cpp:196-236)
 * links rows of the same cluster index whose |r| reaches the threshold, with
 * ones on the diagonal: it is the direct sum of one block per connected
 * component of that graph, and its spectrum is the union of theirs.  Each
 * component is solved on its own, and only when it changed since the
 * previous (higher) threshold.
 */
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kinc_rmt.h"
#include "../../include/kinc_sim.h"

/* ---------------------------------------------------------------- kernels */

/* corr[k][x][y] = corr[k][y][x] = r for every record (dumpRawData) */
__global__ void k_rmt_scatter(const int32_t* __restrict__ x, const int32_t* __restrict__ y,
                              const int8_t* __restrict__ k, const float* __restrict__ r, int64_t n,
                              float* __restrict__ corr, int G)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   float* plane = corr + (size_t)k[i] * (size_t)G * (size_t)G;
   plane[(size_t)x[i] * G + y[i]] = r[i];
   plane[(size_t)y[i] * G + x[i]] = r[i];
}

/* A[i][j] = 1 on the diagonal, else corr[idx_i][idx_j] if |.| >= t, else 0:
 * the pruned block of one cluster index (symmetric, so the row-major fill is
 * also the column-major matrix rocSOLVER reads). */
__global__ void k_rmt_gather(const float* __restrict__ plane, int G, const int32_t* __restrict__ idx,
                             int n, float t, double* __restrict__ A)
{
   const int i = blockIdx.x;
   const int j = blockIdx.y * blockDim.x + threadIdx.x;
   if ( j >= n ) return;
   const float v = plane[(size_t)idx[i] * G + idx[j]];
   A[(size_t)i * n + j] = (i == j) ? 1.0 : (fabsf(v) >= t ? (double)v : 0.0);
}

/* ------------------------------------------------------------------ host */

namespace {

thread_local std::string g_rmt_err;

int set_err(std::string* dst, int code, const std::string& m)
{
   g_rmt_err = m;
   if ( dst ) *dst = m;
   return code;
}

/* Akima spline with GSL's non-periodic end conditions (the interpolation
 * gsl_interp_akima provides, Akima 1970): slopes m_i of the knots, two
 * extrapolated slopes at each end, node slopes weighted by the slope
 * differences, a straight segment where both weights vanish. */
struct Akima
{
   std::vector<double> x, y, b, c, d;

   /* false where GSL refuses the knots (fewer than 5, not increasing) */
   bool init(const std::vector<double>& xs, const std::vector<double>& ys)
   {
      const int n = (int)xs.size();
      if ( n < 5 ) return false;
      for ( int i = 0; i + 1 < n; ++i )
      {
         if ( !(xs[i] < xs[i + 1]) ) return false;
      }
      x = xs;
      y = ys;
      std::vector<double> mv((size_t)n + 3);
      double* m = mv.data() + 2;
      for ( int i = 0; i <= n - 2; ++i ) m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
      m[-2] = 3.0 * m[0] - 2.0 * m[1];
      m[-1] = 2.0 * m[0] - m[1];
      m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
      m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
      b.assign((size_t)n, 0.0);
      c.assign((size_t)n, 0.0);
      d.assign((size_t)n, 0.0);
      for ( int i = 0; i < n - 1; ++i )
      {
         const double w = std::fabs(m[i + 1] - m[i]) + std::fabs(m[i - 1] - m[i - 2]);
         if ( w == 0.0 )
         {
            b[i] = m[i];
            continue;
         }
         const double h = x[i + 1] - x[i];
         const double w_next = std::fabs(m[i + 2] - m[i + 1]) + std::fabs(m[i] - m[i - 1]);
         const double a = std::fabs(m[i - 1] - m[i - 2]) / w;
         const double t_next = (w_next == 0.0)
                                  ? m[i]
                                  : (1.0 - std::fabs(m[i] - m[i - 1]) / w_next) * m[i]
                                       + (std::fabs(m[i] - m[i - 1]) / w_next) * m[i + 1];
         b[i] = (1.0 - a) * m[i - 1] + a * m[i];
         c[i] = (3.0 * m[i] - 2.0 * b[i] - t_next) / h;
         d[i] = (b[i] + t_next - 2.0 * m[i]) / (h * h);
      }
      return true;
   }

   bool inside(double v) const { return v >= x.front() && v <= x.back(); }

   double eval(double v) const
   {
      /* the last knot interval starting at or below v */
      int lo = 0, hi = (int)x.size() - 1;
      while ( hi > lo + 1 )
      {
         const int mid = (lo + hi) / 2;
         if ( x[mid] > v ) hi = mid;
         else lo = mid;
      }
      const double dx = v - x[lo];
      return y[lo] + dx * (b[lo] + dx * (c[lo] + d[lo] * dx));
   }
};

/* RMT::computePaceChiSquare with RMT::unfold (rmt.cpp:320-359,392-439);
 * false where GSL would fail. */
bool pace_chi_square(const std::vector<float>& e, int pace, int bins, float* out)
{
   const int n = (int)e.size();
   const int ns = n / pace;
   std::vector<double> xs((size_t)ns), ys((size_t)ns);
   for ( int i = 0; i < ns; ++i )
   {
      xs[i] = (double)e[(size_t)i * pace];
      ys[i] = (double)(i * pace + 1) / n;
   }
   xs[ns - 1] = e[n - 1];
   ys[ns - 1] = 1.0;
   Akima s;
   if ( !s.init(xs, ys) ) return false;
   std::vector<float> unfolded((size_t)n);
   unfolded[0] = 0.0f;
   unfolded[n - 1] = 1.0f;
   for ( int i = 1; i < n - 1; ++i )
   {
      if ( !s.inside(e[i]) ) return false;
      unfolded[i] = (float)s.eval(e[i]);
   }
   const float hmin = 0.0f, hmax = 3.0f;
   const float width = (hmax - hmin) / bins;
   std::vector<float> hist((size_t)bins, 0.0f);
   for ( int i = 0; i < n - 1; ++i )
   {
      const float spacing = (unfolded[i + 1] - unfolded[i]) * n;
      if ( hmin <= spacing && spacing < hmax ) ++hist[(int)((spacing - hmin) / width)];
   }
   float chi = 0.0f;
   for ( int i = 0; i < bins; ++i )
   {
      const float O = hist[i];
      const float E = (expf(-i * width) - expf(-(i + 1) * width)) * n;
      chi += (O - E) * (O - E) / E;
   }
   *out = chi;
   return true;
}

} // namespace

struct kinc_rmt
{
   kinc_rmt_params p;
   int G = 0, K = 0, device = 0;
   hipStream_t stream = nullptr;
   rocblas_handle blas = nullptr;
   float* d_corr = nullptr;
   int32_t* d_idx = nullptr;
   double* d_A = nullptr;
   double* d_D = nullptr;
   double* d_E = nullptr;
   rocblas_int* d_info = nullptr;
   std::string err;

   /* Incremental state of a sweep.  Nodes are (gene, cluster) rows, id
    * g*K + k; an edge is a record with |r| >= t.  A node is in the pruned
    * matrix exactly when one of its edges is (its maximum |r| reaches t,
    * computeMaximums/computePruneMatrix), and the pruned matrix is the
    * direct sum of the connected components of this graph, so its spectrum
    * is the union of theirs.  As t decreases, edges are only added: the
    * components are kept in a union-find and a component's eigenvalues are
    * recomputed only when it gained an edge. */
   struct Edge
   {
      int32_t a, b;
      float r;
   };
   std::vector<Edge> edges;                 /* by decreasing |r| */
   size_t next = 0;
   float last_t = INFINITY;
   std::vector<int32_t> parent;             /* -1: not in the pruned matrix */
   std::vector<std::vector<int32_t>> members;
   std::vector<std::vector<double>> eig;
   std::vector<float> pair_r;               /* the edge of a two-node component */
   std::vector<int64_t> stamp;
   std::vector<int32_t> dirty;
   int64_t step = 0;
};

#define RMT_HIP(h, call)                                                                     \
   do                                                                                        \
   {                                                                                         \
      hipError_t e_ = (call);                                                                \
      if ( e_ != hipSuccess )                                                                \
         return set_err((h) ? &(h)->err : nullptr, KINC_E_DEVICE,                               \
                        std::string(#call) + ": " + hipGetErrorString(e_));                  \
   } while ( 0 )

extern "C" void kinc_rmt_default_params(kinc_rmt_params* p)
{
   if ( !p ) return;
   memset(p, 0, sizeof(*p));
   p->threshold_start = 0.99f;
   p->threshold_step = 0.001f;
   p->threshold_stop = 0.5f;
   p->chi_square1 = 99.607f;
   p->chi_square2 = 200.0f;
   p->min_eigenvalues = 50;
   p->min_pace = 10;
   p->max_pace = 40;
   p->bins = 60;
}

extern "C" void kinc_rmt_destroy(kinc_rmt* h)
{
   if ( !h ) return;
   (void)hipSetDevice(h->device);
   if ( h->blas ) rocblas_destroy_handle(h->blas);
   (void)hipFree(h->d_corr);
   (void)hipFree(h->d_idx);
   (void)hipFree(h->d_A);
   (void)hipFree(h->d_D);
   (void)hipFree(h->d_E);
   (void)hipFree(h->d_info);
   if ( h->stream ) (void)hipStreamDestroy(h->stream);
   delete h;
}

extern "C" const char* kinc_rmt_last_error(const kinc_rmt* h)
{
   return h ? h->err.c_str() : g_rmt_err.c_str();
}

static int check_params(const kinc_rmt_params* p, std::string* err)
{
   /* RMT::initialize (rmt.cpp:129-153) */
   if ( p->threshold_start <= p->threshold_stop )
      return set_err(err, KINC_E_INVALID, "Starting threshold must be greater than stopping threshold.");
   if ( p->min_pace >= p->max_pace )
      return set_err(err, KINC_E_INVALID, "Minimum unfolding pace must be less than maximum unfolding pace.");
   if ( !(p->threshold_step > 0) || p->min_pace < 1 || p->bins < 1 || p->min_eigenvalues < 0 )
      return set_err(err, KINC_E_INVALID, "invalid RMT parameters");
   return KINC_OK;
}

extern "C" int kinc_rmt_create(const kinc_rmt_params* params, int32_t G, int32_t K, const int32_t* x,
                               const int32_t* y, const int8_t* cluster, const float* corr,
                               int64_t n, int32_t device, kinc_rmt** out)
{
   if ( !params || !out || G < 2 || K < 1 || K > 64 || n < 0 || (n > 0 && (!x || !y || !cluster || !corr)) )
      return set_err(nullptr, KINC_E_INVALID, "invalid RMT arguments");
   int rc = check_params(params, nullptr);
   if ( rc ) return rc;
   /* records inside the matrix, in the strictly increasing order of a CMX file */
   int64_t last = -1;
   for ( int64_t i = 0; i < n; ++i )
   {
      if ( !(0 <= y[i] && y[i] < x[i] && x[i] < G) || cluster[i] < 0 || cluster[i] >= K )
         return set_err(nullptr, KINC_E_INVALID, "record " + std::to_string(i) + " outside the matrix");
      const int64_t indent = ((int64_t)x[i] * (x[i] - 1) / 2 + y[i]) * 64 + cluster[i];
      if ( indent <= last )
         return set_err(nullptr, KINC_E_INVALID, "records must be in strictly increasing pair order");
      last = indent;
   }
   int ndev = 0;
   if ( hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev )
      return set_err(nullptr, KINC_E_DEVICE, "no such HIP device");

   kinc_rmt* h = new kinc_rmt();
   h->p = *params;
   h->G = G;
   h->K = K;
   h->device = device;
   auto bail = [&](int code) {
      std::string m = h->err;
      kinc_rmt_destroy(h);
      return set_err(nullptr, code, m);
   };
#define CREATE_HIP(call)                                                                     \
   do                                                                                        \
   {                                                                                         \
      hipError_t e_ = (call);                                                                \
      if ( e_ != hipSuccess )                                                                \
      {                                                                                      \
         h->err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
         return bail(KINC_E_DEVICE);                                                            \
      }                                                                                      \
   } while ( 0 )
   CREATE_HIP(hipSetDevice(device));
   CREATE_HIP(hipStreamCreate(&h->stream));
   if ( rocblas_create_handle(&h->blas) != rocblas_status_success
        || rocblas_set_stream(h->blas, h->stream) != rocblas_status_success )
   {
      h->err = "rocblas handle";
      return bail(KINC_E_DEVICE);
   }
   const size_t plane = (size_t)G * (size_t)G;
   size_t free_b = 0, total_b = 0;
   CREATE_HIP(hipMemGetInfo(&free_b, &total_b));
   const size_t need = sizeof(float) * plane * (size_t)K + sizeof(double) * plane + ((size_t)64 << 20);
   if ( need > free_b )
   {
      h->err = "the correlation matrix (" + std::to_string(need >> 20) + " MiB with workspace) does not fit in device memory";
      return bail(KINC_E_NOMEM);
   }
   CREATE_HIP(hipMalloc(&h->d_corr, sizeof(float) * plane * (size_t)K));
   CREATE_HIP(hipMemsetAsync(h->d_corr, 0, sizeof(float) * plane * (size_t)K, h->stream));
   CREATE_HIP(hipMalloc(&h->d_idx, sizeof(int32_t) * (size_t)G));
   CREATE_HIP(hipMalloc(&h->d_A, sizeof(double) * plane));
   CREATE_HIP(hipMalloc(&h->d_D, sizeof(double) * (size_t)G));
   CREATE_HIP(hipMalloc(&h->d_E, sizeof(double) * (size_t)G));
   CREATE_HIP(hipMalloc(&h->d_info, sizeof(rocblas_int)));

   /* records in chunks through one staging set */
   const int64_t chunk = 1 << 22;
   if ( n > 0 )
   {
      int32_t *dx = nullptr, *dy = nullptr;
      int8_t* dk = nullptr;
      float* dr = nullptr;
      const int64_t c0 = std::min(n, chunk);
      CREATE_HIP(hipMalloc(&dx, sizeof(int32_t) * c0));
      CREATE_HIP(hipMalloc(&dy, sizeof(int32_t) * c0));
      CREATE_HIP(hipMalloc(&dk, (size_t)c0));
      CREATE_HIP(hipMalloc(&dr, sizeof(float) * c0));
      hipError_t e = hipSuccess;
      for ( int64_t s = 0; s < n && e == hipSuccess; s += chunk )
      {
         const int64_t m = std::min(chunk, n - s);
         e = hipMemcpyAsync(dx, x + s, sizeof(int32_t) * m, hipMemcpyHostToDevice, h->stream);
         if ( e == hipSuccess ) e = hipMemcpyAsync(dy, y + s, sizeof(int32_t) * m, hipMemcpyHostToDevice, h->stream);
         if ( e == hipSuccess ) e = hipMemcpyAsync(dk, cluster + s, (size_t)m, hipMemcpyHostToDevice, h->stream);
         if ( e == hipSuccess ) e = hipMemcpyAsync(dr, corr + s, sizeof(float) * m, hipMemcpyHostToDevice, h->stream);
         if ( e == hipSuccess )
         {
            k_rmt_scatter<<<(unsigned)((m + 255) / 256), 256, 0, h->stream>>>(dx, dy, dk, dr, m, h->d_corr, G);
            e = hipGetLastError();
         }
         /* the staging set is reused: finish this chunk first */
         if ( e == hipSuccess ) e = hipStreamSynchronize(h->stream);
      }
      (void)hipFree(dx);
      (void)hipFree(dy);
      (void)hipFree(dk);
      (void)hipFree(dr);
      CREATE_HIP(e);
   }
   CREATE_HIP(hipStreamSynchronize(h->stream));
#undef CREATE_HIP
   h->edges.reserve((size_t)n);
   for ( int64_t i = 0; i < n; ++i )
   {
      if ( corr[i] != 0.0f ) h->edges.push_back({x[i] * K + cluster[i], y[i] * K + cluster[i], corr[i]});
   }
   std::stable_sort(h->edges.begin(), h->edges.end(),
                    [](const kinc_rmt::Edge& a, const kinc_rmt::Edge& b) { return std::fabs(a.r) > std::fabs(b.r); });
   const size_t nodes = (size_t)G * (size_t)K;
   h->parent.assign(nodes, -1);
   h->members.resize(nodes);
   h->eig.resize(nodes);
   h->pair_r.assign(nodes, 0.0f);
   h->stamp.assign(nodes, -1);
   *out = h;
   return KINC_OK;
}

static int32_t find_root(kinc_rmt* h, int32_t v)
{
   while ( h->parent[v] != v )
   {
      h->parent[v] = h->parent[h->parent[v]];
      v = h->parent[v];
   }
   return v;
}

static void activate(kinc_rmt* h, int32_t v)
{
   if ( h->parent[v] >= 0 ) return;
   h->parent[v] = v;
   h->members[v].assign(1, v);
   h->eig[v].clear();
   h->dirty.push_back(v);
}

static void reset_sweep(kinc_rmt* h)
{
   for ( size_t v = 0; v < h->parent.size(); ++v )
   {
      if ( h->parent[v] < 0 ) continue;
      h->parent[v] = -1;
      h->members[v].clear();
      h->eig[v].clear();
   }
   h->next = 0;
   h->dirty.clear();
}

/* Eigenvalues of one component: closed forms for one node (1) and two nodes
 * (1 - r, 1 + r, exact in double), else its block gathered on the device
 * and solved by rocSOLVER. */
static int solve_component(kinc_rmt* h, int32_t root, float t)
{
   std::vector<double>& w = h->eig[root];
   const std::vector<int32_t>& mem = h->members[root];
   const int c = (int)mem.size();
   if ( c == 1 )
   {
      w.assign(1, 1.0);
      return KINC_OK;
   }
   if ( c == 2 )
   {
      const double r = std::fabs((double)h->pair_r[root]);
      w = {1.0 - r, 1.0 + r};
      return KINC_OK;
   }
   const int G = h->G, K = h->K;
   const int k = mem[0] % K;
   std::vector<int32_t> genes((size_t)c);
   for ( int i = 0; i < c; ++i ) genes[i] = mem[i] / K;
   std::sort(genes.begin(), genes.end());
   RMT_HIP(h, hipMemcpyAsync(h->d_idx, genes.data(), sizeof(int32_t) * c, hipMemcpyHostToDevice, h->stream));
   dim3 grid((unsigned)c, (unsigned)((c + 255) / 256));
   k_rmt_gather<<<grid, 256, 0, h->stream>>>(h->d_corr + (size_t)k * G * G, G, h->d_idx, c, t, h->d_A);
   RMT_HIP(h, hipGetLastError());
   rocblas_status st = rocsolver_dsyevd(h->blas, rocblas_evect_none, rocblas_fill_lower, c, h->d_A, c,
                                        h->d_D, h->d_E, h->d_info);
   if ( st != rocblas_status_success )
      return set_err(&h->err, KINC_E_DEVICE, "rocsolver_dsyevd failed: " + std::to_string((int)st));
   w.resize((size_t)c);
   rocblas_int info = 0;
   RMT_HIP(h, hipMemcpyAsync(w.data(), h->d_D, sizeof(double) * c, hipMemcpyDeviceToHost, h->stream));
   RMT_HIP(h, hipMemcpyAsync(&info, h->d_info, sizeof(info), hipMemcpyDeviceToHost, h->stream));
   RMT_HIP(h, hipStreamSynchronize(h->stream));
   if ( info != 0 )
      return set_err(&h->err, KINC_E_DEVICE, "dsyevd did not converge (info " + std::to_string(info) + ")");
   return KINC_OK;
}

/* Eigenvalues of the matrix pruned at t, |.|-ascending (double). */
static int eigenvalues(kinc_rmt* h, float t, std::vector<double>& w)
{
   if ( t > h->last_t ) reset_sweep(h);
   h->last_t = t;
   ++h->step;
   /* add the edges that reach t; mark the components they touch */
   while ( h->next < h->edges.size() && std::fabs(h->edges[h->next].r) >= t )
   {
      const kinc_rmt::Edge& e = h->edges[h->next++];
      activate(h, e.a);
      activate(h, e.b);
      int32_t ra = find_root(h, e.a), rb = find_root(h, e.b);
      if ( ra != rb )
      {
         if ( h->members[ra].size() < h->members[rb].size() ) std::swap(ra, rb);
         h->members[ra].insert(h->members[ra].end(), h->members[rb].begin(), h->members[rb].end());
         std::vector<int32_t>().swap(h->members[rb]);
         std::vector<double>().swap(h->eig[rb]);
         h->parent[rb] = ra;
      }
      if ( h->members[ra].size() == 2 ) h->pair_r[ra] = e.r;
      h->dirty.push_back(ra);
   }
   for ( int32_t v : h->dirty )
   {
      const int32_t r = find_root(h, v);
      if ( h->stamp[r] == h->step ) continue;
      h->stamp[r] = h->step;
      int rc = solve_component(h, r, t);
      if ( rc ) return rc;
   }
   h->dirty.clear();
   w.clear();
   for ( size_t v = 0; v < h->parent.size(); ++v )
   {
      if ( h->parent[v] == (int32_t)v ) w.insert(w.end(), h->eig[v].begin(), h->eig[v].end());
   }
   /* gsl_eigen_symmv_sort(GSL_EIGEN_SORT_ABS_ASC) */
   std::stable_sort(w.begin(), w.end(), [](double a, double b) { return std::fabs(a) < std::fabs(b); });
   return KINC_OK;
}

extern "C" int kinc_rmt_eigenvalues(kinc_rmt* h, float threshold, float* eigens, int64_t cap, int64_t* size)
{
   if ( !h || !size ) return set_err(nullptr, KINC_E_INVALID, "invalid arguments");
   RMT_HIP(h, hipSetDevice(h->device));
   std::vector<double> w;
   int rc = eigenvalues(h, threshold, w);
   if ( rc ) return rc;
   *size = (int64_t)w.size();
   if ( eigens )
   {
      if ( cap < (int64_t)w.size() ) return set_err(&h->err, KINC_E_INVALID, "eigenvalue buffer too small");
      for ( size_t i = 0; i < w.size(); ++i ) eigens[i] = (float)w[i];
   }
   return KINC_OK;
}

extern "C" float kinc_rmt_chi_square(const kinc_rmt_params* p, const float* eigens, int64_t n)
{
   /* RMT::degenerate (rmt.cpp:366-383): values 1..n-1, dropping those
    * within 1e-6 of the last kept one */
   const float EPSILON = 1e-6f;
   std::vector<float> u;
   for ( int64_t i = 1; i < n; ++i )
   {
      if ( u.empty() || std::fabs(eigens[i] - u.back()) > EPSILON ) u.push_back(eigens[i]);
   }
   if ( (int64_t)u.size() < p->min_eigenvalues ) return -1;
   float chi = 0.0f;
   int count = 0;
   for ( int pace = p->min_pace; pace <= p->max_pace; ++pace )
   {
      if ( (int)u.size() / pace < 5 ) break;
      float c = 0.0f;
      if ( !pace_chi_square(u, pace, p->bins, &c) ) return -2;
      chi += c;
      ++count;
   }
   chi /= count;
   return chi;
}

extern "C" int kinc_rmt_run(kinc_rmt* h, float* log_t, int32_t* log_size, float* log_chi, int64_t cap,
                            int64_t* steps, float* final_threshold)
{
   if ( !h || !steps || !final_threshold ) return set_err(nullptr, KINC_E_INVALID, "invalid arguments");
   RMT_HIP(h, hipSetDevice(h->device));
   const kinc_rmt_params& p = h->p;
   float finalThreshold = 0, finalChi = INFINITY, maxChi = -INFINITY;
   float threshold = p.threshold_start;
   int64_t s = 0;
   std::vector<double> w;
   std::vector<float> wf;
   *steps = 0;
   *final_threshold = 0;
   while ( maxChi < p.chi_square2 )
   {
      int rc = eigenvalues(h, threshold, w);
      if ( rc ) return rc;
      const int64_t size = (int64_t)w.size();
      float chi = -1;
      if ( size > 0 )
      {
         wf.resize(w.size());
         for ( size_t i = 0; i < w.size(); ++i ) wf[i] = (float)w[i];
         chi = kinc_rmt_chi_square(&p, wf.data(), size);
      }
      if ( chi == -2 )
      {
         /* the reference aborts inside this step: nothing of it is logged */
         return set_err(&h->err, KINC_E_INVALID,
                        "spline knots not strictly increasing at threshold " + std::to_string(threshold)
                           + " (negative eigenvalues in |value| order; GSL would abort)");
      }
      if ( s < cap )
      {
         if ( log_t ) log_t[s] = threshold;
         if ( log_size ) log_size[s] = (int32_t)size;
         if ( log_chi ) log_chi[s] = chi;
      }
      *steps = ++s;
      if ( chi != -1 )
      {
         if ( chi < p.chi_square1 )
         {
            finalChi = chi;
            finalThreshold = threshold;
         }
         if ( finalChi < p.chi_square1 && chi > finalChi ) maxChi = chi;
      }
      threshold -= p.threshold_step;
      if ( threshold < p.threshold_stop )
      {
         return set_err(&h->err, KINC_E_INVALID,
                        "RMT Threshold Error: Could not find non-random threshold above stopping threshold.");
      }
   }
   *final_threshold = finalThreshold;
   return KINC_OK;
}

/* ------------------------------------------------------------ CMX input */

namespace {

uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

/* The records of a CMX file in this library's envelope (kinc_formats.h):
 * the first correlation of each record. */
int read_cmx(const char* path, int32_t* G, int32_t* K, std::vector<int32_t>& x, std::vector<int32_t>& y,
             std::vector<int8_t>& k, std::vector<float>& r)
{
   FILE* f = fopen(path, "rb");
   if ( !f ) return set_err(nullptr, KINC_E_INVALID, std::string(path) + ": cannot open");
   uint8_t env[64];
   int rc = KINC_OK;
   if ( fread(env, 1, 64, f) != 64 || memcmp(env, "KINCAMD1", 8) != 0 || rd32(env + 8) != 3 )
      rc = set_err(nullptr, KINC_E_INVALID, std::string(path) + ": not a CMX file");
   std::vector<uint8_t> payload;
   if ( rc == KINC_OK )
   {
      const uint64_t off = rd64(env + 16), size = rd64(env + 24);
      payload.resize((size_t)size);
      if ( fseek(f, (long)off, SEEK_SET) != 0 || fread(payload.data(), 1, payload.size(), f) != payload.size() )
         rc = set_err(nullptr, KINC_E_INVALID, std::string(path) + ": short file");
   }
   fclose(f);
   if ( rc ) return rc;
   if ( payload.size() < 30 ) return set_err(nullptr, KINC_E_INVALID, std::string(path) + ": short header");
   const uint8_t* p = payload.data();
   *G = (int32_t)rd32(p);
   *K = (int32_t)rd32(p + 4);
   const int32_t data_size = (int32_t)rd32(p + 8);
   const int64_t clusters = (int64_t)rd64(p + 20);
   const int off = p[28] | (p[29] << 8);
   const size_t rec = 9 + (size_t)data_size;
   const size_t base = 30 + (size_t)off;
   if ( data_size < 4 || base + rec * (size_t)clusters > payload.size() )
      return set_err(nullptr, KINC_E_INVALID, std::string(path) + ": truncated records");
   x.resize((size_t)clusters);
   y.resize((size_t)clusters);
   k.resize((size_t)clusters);
   r.resize((size_t)clusters);
   for ( int64_t i = 0; i < clusters; ++i )
   {
      const uint8_t* q = p + base + rec * (size_t)i;
      x[i] = (int32_t)rd32(q);
      y[i] = (int32_t)rd32(q + 4);
      k[i] = (int8_t)q[8];
      memcpy(&r[i], q + 9, 4);
   }
   return KINC_OK;
}

} // namespace

extern "C" int kinc_rmt_run_file(const kinc_rmt_params* params, const char* cmx_path, const char* log_path,
                                 int32_t device, float* final_threshold)
{
   if ( !params || !cmx_path || !log_path || !final_threshold )
      return set_err(nullptr, KINC_E_INVALID, "Did not get valid input or logfile arguments.");
   int32_t G = 0, K = 0;
   std::vector<int32_t> x, y;
   std::vector<int8_t> k;
   std::vector<float> r;
   int rc = read_cmx(cmx_path, &G, &K, x, y, k, r);
   if ( rc ) return rc;
   kinc_rmt* h = nullptr;
   rc = kinc_rmt_create(params, G, K, x.data(), y.data(), k.data(), r.data(), (int64_t)x.size(), device, &h);
   if ( rc ) return rc;
   const int64_t cap = (int64_t)((params->threshold_start - params->threshold_stop) / params->threshold_step) + 16;
   std::vector<float> lt((size_t)cap), lc((size_t)cap);
   std::vector<int32_t> ls((size_t)cap);
   int64_t steps = 0;
   float t = 0;
   rc = kinc_rmt_run(h, lt.data(), ls.data(), lc.data(), cap, &steps, &t);
   const std::string msg = h->err;
   kinc_rmt_destroy(h);
   FILE* f = fopen(log_path, "w");
   if ( !f ) return set_err(nullptr, KINC_E_INVALID, std::string(log_path) + ": cannot open for writing");
   for ( int64_t s = 0; s < std::min(steps, cap); ++s )
      fprintf(f, "%g\t%d\t%g\n", (double)lt[s], ls[s], (double)lc[s]);
   if ( rc == KINC_OK ) fprintf(f, "%g\n", (double)t);
   fclose(f);
   if ( rc ) return set_err(nullptr, rc, msg);
   *final_threshold = t;
   return KINC_OK;
}
