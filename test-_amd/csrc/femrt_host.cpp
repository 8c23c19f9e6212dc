// femrt_host.cpp -- host-side algorithms of the library: GLL nodes, modal
// fitting, the structured curved-hex generator, exact Bernstein bounding
// boxes and the binned-SAH BVH builder.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "femrt_internal.hpp"

namespace femrt {

// Runs body(e0, e1) on contiguous chunks of [0, n), one host thread per chunk
// (at most 16).  Used for per-element work whose result does not depend on
// the split.
template <typename F>
static void parallel_chunks(int n, F body)
{
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 512) nt = 1;
  if (nt <= 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back(body, (int)((long long)n * t / nt), (int)((long long)n * (t + 1) / nt));
  for (std::thread& x : th) x.join();
}

// ---------------------------------------------------------------- 1-D tools

// Modified_A basis value (same definition as basis1d in femrt_math.hpp, in
// double, used by the host fitting code).
double modA(int p, double x)
{
  if (p == 0) return 0.5 * (1.0 - x);
  if (p == 1) return 0.5 * (1.0 + x);
  const double b = 0.25 * (1.0 - x * x);
  double pm2 = 1.0, pm1 = 2.0 * x;
  if (p == 2) return b;
  if (p == 3) return b * pm1;
  double pk = pm1;
  for (int q = 4; q <= p; ++q) {
    const int k = q - 2;
    const double A = (2.0 * k + 1.0) * (k + 1.0) / (k * (k + 2.0));
    const double B = (k + 1.0) / (k + 2.0);
    pk = A * x * pm1 - B * pm2;
    pm2 = pm1;
    pm1 = pk;
  }
  return b * pk;
}

// Jacobi polynomial P_n^(a,b)(x) by the textbook three-term recurrence
// (Karniadakis & Sherwin, Appendix A), in double.
static double jacobi(int n, double a, double b, double x)
{
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = 0.5 * ((a + b + 2.0) * x + (a - b));
  for (int k = 2; k <= n; ++k) {
    const double c = 2.0 * k + a + b;
    const double p2 = ((c - 1.0) * (c * (c - 2.0) * x + a * a - b * b) * p1 -
                       2.0 * (k + a - 1.0) * (k + b - 1.0) * c * p0) /
                      (2.0 * k * (k + a + b) * (c - 2.0));
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// The prism's collapsed-direction basis as the reference defines it
// (ModifiedB, fem/ModalBasis.cuh:72-82): Modified_A(k) for i = 0,
// ((1-x)/2)^i for k = 0, else ((1-x)/2)^i (1+x)/2 P_{k-1}^(2i-1,1)(x).  A
// polynomial of degree i + k in x.
double modB(int i, int k, double x)
{
  if (i == 0) return modA(k, x);
  const double w = std::pow(0.5 * (1.0 - x), i);
  if (k == 0) return w;
  return w * 0.5 * (1.0 + x) * jacobi(k - 1, 2.0 * i - 1.0, 1.0, x);
}

// Number of prism coefficients for modes (mx, my, mz): i < mx slowest, then
// j < my, then k < mz - i fastest (EvaluatePrismModal,
// fem/PrismElement.cu:16-37).
int prism_n_coeffs(const int* m)
{
  int c = 0;
  for (int i = 0; i < m[0]; ++i) c += m[1] * std::max(m[2] - i, 0);
  return c;
}

// Modified_B(i, k) in the Modified_A basis of n modes: b[(i * mc + k) * n + q]
// for i < mi, k < mc - i, by interpolation at n GLL nodes (exact: degree
// i + k <= mc - 1 <= n - 1).
static std::vector<double> modB_in_modA(int mi, int mc, int n)
{
  const std::vector<double> nodes = gll_nodes(n);
  std::vector<double> inv;
  modal_fit_1d_matrix(n, nodes, inv);
  std::vector<double> b((size_t)mi * mc * n, 0.0);
  for (int i = 0; i < mi; ++i)
    for (int k = 0; k + i < mc; ++k)
      for (int q = 0; q < n; ++q) {
        double s = 0.0;
        for (int p = 0; p < n; ++p) s += inv[(size_t)q * n + p] * modB(i, k, nodes[p]);
        b[((size_t)i * mc + k) * n + q] = s;
      }
  return b;
}

// Triangle coefficients (per element `comps` blocks: i < mx slowest, j <
// my - i, Modified_A(i, r) Modified_B(i, j, s); EvaluateTriModal,
// fem/TriElement.cu:46-63) as quad coefficients with n modes per direction,
// j n + i, in double.  Needs n >= mx, my.
void tri_to_quad(int n_elem, const int* m, int n, const float* coeffs, int comps, double* out)
{
  const std::vector<double> b = modB_in_modA(m[0], m[1], n);
  int nc = 0;
  for (int i = 0; i < m[0]; ++i) nc += std::max(m[1] - i, 0);
  const size_t n2 = (size_t)n * n;
  parallel_chunks(n_elem * comps, [&](int e0, int e1) {
    for (int ec = e0; ec < e1; ++ec) {
      const float* c = coeffs + (size_t)ec * nc;
      double* h = out + (size_t)ec * n2;
      std::fill(h, h + n2, 0.0);
      int idx = 0;
      for (int i = 0; i < m[0]; ++i)
        for (int j = 0; j + i < m[1]; ++j, ++idx)
          for (int q = 0; q < n; ++q) h[(size_t)q * n + i] += (double)c[idx] * b[((size_t)i * m[1] + j) * n + q];
    }
  });
}

// Prism coefficients (per element `comps` blocks in the reference's layout)
// as hex coefficients with n modes per direction, (k n + j) n + i, in double:
// sum_{i,j,k} c_ijk A_i(r) A_j(s) B_ik(t) = sum_{i,j,m} H_mji A_i(r) A_j(s) A_m(t)
// with B_ik(t) = sum_m b_ikm A_m(t), exact since B_ik has degree
// i + k <= mz - 1 <= n - 1 (b by interpolation at n GLL nodes).  Needs
// n >= mx, my, mz.
void prism_to_hex(int n_elem, const int* m, int n, const float* coeffs, int comps, double* out)
{
  const std::vector<double> b = modB_in_modA(m[0], m[2], n);   // b[(i * mz + k) * n + q]
  const int nc = prism_n_coeffs(m);
  const size_t n3 = (size_t)n * n * n;
  parallel_chunks(n_elem * comps, [&](int e0, int e1) {
    for (int ec = e0; ec < e1; ++ec) {
      const float* c = coeffs + (size_t)ec * nc;
      double* h = out + (size_t)ec * n3;
      std::fill(h, h + n3, 0.0);
      int idx = 0;
      for (int i = 0; i < m[0]; ++i)
        for (int j = 0; j < m[1]; ++j)
          for (int k = 0; k + i < m[2]; ++k, ++idx)
            for (int q = 0; q < n; ++q) h[((size_t)q * n + j) * n + i] += (double)c[idx] * b[((size_t)i * m[2] + k) * n + q];
    }
  });
}

static double legendre(int n, double x, double* pnm1)
{
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    if (pnm1) *pnm1 = 0.0;
    return 1.0;
  }
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  if (pnm1) *pnm1 = p0;
  return p1;
}

// Gauss-Lobatto-Legendre nodes of n points on [-1,1], ascending: the roots of
// (1-x^2) P'_{n-1}(x).  Newton iteration on x P_N - P_{N-1} = 0 started from
// the Chebyshev-Gauss-Lobatto points (Canuto, Hussaini, Quarteroni & Zang,
// Spectral Methods, 2.3).  The reference computes the same nodes
// (xml_fld_math.cpp:60-112) and stores them as float.
std::vector<double> gll_nodes(int n)
{
  std::vector<double> x(n);
  if (n == 1) {
    x[0] = 0.0;
    return x;
  }
  const int N = n - 1;
  for (int i = 0; i <= N; ++i) {
    double xi = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double pm1;
      const double p = legendre(N, xi, &pm1);
      // f = x P_N - P_{N-1};  f' = (N+1) P_N
      const double dx = (xi * p - pm1) / ((N + 1) * p);
      xi -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    x[i] = xi;
  }
  x[0] = -1.0;
  x[N] = 1.0;
  for (int i = 0; i < n / 2; ++i) {  // exact antisymmetry
    const double s = 0.5 * (x[N - i] - x[i]);
    x[i] = -s;
    x[N - i] = s;
  }
  if (n % 2) x[n / 2] = 0.0;
  return x;
}

// Gauss-Jordan inverse with partial pivoting (double).
static bool invert(std::vector<double> a, int n, std::vector<double>& inv)
{
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (std::fabs(a[(size_t)piv * n + c]) < 1e-300) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double ip = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= ip;
      inv[(size_t)c * n + k] *= ip;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// inv = V^{-1}, V[i][a] = phi_a(node_i): modal coefficients from nodal values
// (the 1-D factor of TensorModalFit3D, xml_fld_math.cpp:379-483).
bool modal_fit_1d_matrix(int n, const std::vector<double>& nodes, std::vector<double>& inv)
{
  std::vector<double> V((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < n; ++a) V[(size_t)i * n + a] = modA(a, nodes[i]);
  return invert(V, n, inv);
}

// Lagrange interpolation matrix from nodes src (n) to nodes dst (m): m x n.
std::vector<double> lagrange_matrix(const std::vector<double>& src,
                                    const std::vector<double>& dst)
{
  const int n = (int)src.size(), m = (int)dst.size();
  std::vector<double> L((size_t)m * n);
  for (int i = 0; i < m; ++i)
    for (int a = 0; a < n; ++a) {
      double v = 1.0;
      for (int b = 0; b < n; ++b)
        if (b != a) v *= (dst[i] - src[b]) / (src[a] - src[b]);
      L[(size_t)i * n + a] = v;
    }
  return L;
}

// Tensor modal fit of nodal values on an n^3 grid (index (k*n+j)*n+i) to
// modal coefficients with the same ordering.
void tensor_fit(int n, const std::vector<double>& inv, const double* nodal, double* out)
{
  std::vector<double> t1((size_t)n * n * n), t2((size_t)n * n * n);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < n; ++a) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += inv[(size_t)a * n + i] * nodal[((size_t)k * n + j) * n + i];
        t1[((size_t)k * n + j) * n + a] = s;
      }
  for (int k = 0; k < n; ++k)
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += inv[(size_t)b * n + j] * t1[((size_t)k * n + j) * n + a];
        t2[((size_t)k * n + b) * n + a] = s;
      }
  for (int c = 0; c < n; ++c)
    for (int b = 0; b < n; ++b)
      for (int a = 0; a < n; ++a) {
        double s = 0;
        for (int k = 0; k < n; ++k) s += inv[(size_t)c * n + k] * t2[((size_t)k * n + b) * n + a];
        out[((size_t)c * n + b) * n + a] = s;
      }
}

// ---------------------------------------------------------- face adjacency

// Device form of the face adjacency: per (element, local face) the neighbour
// element and (neighbour face | orientation << 3), where the orientation o
// maps the face coordinates (u, v) of this side to the neighbour's:
// swap when o & 1, then negate u' when o & 2 and v' when o & 4.  It is found
// from the corner positions (corner c of face f sits at (u, v) = (-1,-1),
// (1,-1), (1,1), (-1,1) for c = 0..3).  A pair whose corners match under no
// orientation is dropped (the walk then finds the next element through the
// BVH).
void encode_adjacency(int n_elem, const float* vertices, const int* adjacency, int* out)
{
  static const int fc[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                               {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
  static const int uv[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  for (int e = 0; e < n_elem; ++e)
    for (int f = 0; f < 6; ++f) {
      const int ne = adjacency[((size_t)e * 6 + f) * 2], nf = adjacency[((size_t)e * 6 + f) * 2 + 1];
      out[((size_t)e * 6 + f) * 2] = -1;
      out[((size_t)e * 6 + f) * 2 + 1] = -1;
      if (ne < 0 || ne >= n_elem || nf < 0 || nf > 5) continue;
      const float* P = vertices + (size_t)e * 24;
      const float* Q = vertices + (size_t)ne * 24;
      double scale = 0.0;
      for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 3; ++k)
          scale = std::max(scale, std::fabs((double)P[3 * fc[f][c] + k] - P[3 * fc[f][0] + k]));
      for (int o = 0; o < 8; ++o) {
        bool all = true;
        for (int c = 0; c < 4 && all; ++c) {
          int u = uv[c][0], v = uv[c][1];
          if (o & 1) std::swap(u, v);
          if (o & 2) u = -u;
          if (o & 4) v = -v;
          int c2 = 0;
          while (uv[c2][0] != u || uv[c2][1] != v) ++c2;
          for (int k = 0; k < 3; ++k)
            if (std::fabs((double)P[3 * fc[f][c] + k] - Q[3 * fc[nf][c2] + k]) > 1e-5 * scale) all = false;
        }
        if (all) {
          out[((size_t)e * 6 + f) * 2] = ne;
          out[((size_t)e * 6 + f) * 2 + 1] = nf | (o << 3);
          break;
        }
      }
    }
}

// ------------------------------------------------------------ random numbers

static inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Standard normal keyed by a tuple: order-independent counter-based draw
// (splitmix64 finaliser + Box-Muller).
static double normal_keyed(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
  uint64_t h = mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
  const uint64_t h2 = mix64(h ^ 0xD1B54A32D192ED03ull);
  const double u1 = ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// ------------------------------------------------------ structured generator

// Local faces in Nektar order with the corner vertices that span (u, v):
// grid (0,0)=c0, (max,0)=c1, (max,max)=c2, (0,max)=c3
// (xml_fld_geometry_builder.cpp:85-97).
static const int kFaceCorners[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                       {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};

int gen_structured_hex(const femrt_gen_params* gp, float* vertices, float* geom, float* field,
                       int* adjacency)
{
  const int N = gp->n, q = gp->q, ng = gp->q + 1, nf = gp->p + 1;
  if (N < 1 || q < 1 || gp->p < 1 || ng > 16 || nf > 16) return FEMRT_EINVAL;
  const double h = 1.0 / N;
  const std::vector<double> gll = gll_nodes(ng);
  std::vector<double> even(ng);
  for (int a = 0; a < ng; ++a) even[a] = -1.0 + 2.0 * a / q;
  const std::vector<double> L = lagrange_matrix(even, gll);  // ng x ng
  std::vector<double> inv;
  modal_fit_1d_matrix(ng, gll, inv);
  const size_t n3 = (size_t)ng * ng * ng, f3 = (size_t)nf * nf * nf;

  // Integer lattice of geometry nodes: a face node of element (i,j,k) at local
  // evenly spaced position (a,b) has lattice coordinates (i*q+.., ...), unique
  // across elements, which keys its displacement.
  auto displacement = [&](int axis, const long lat[3], const double w[3]) -> double {
    if (gp->pert == 1) {
      const int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
      return gp->amp * std::sin(2.0 * M_PI * w[a0]) * std::sin(2.0 * M_PI * w[a1]);
    }
    if (gp->pert == 2 || gp->pert == 3) {
      // seeded bubble per face: amp h xi_f (1-u^2)(1-v^2), xi_f ~ N(0,1)
      // clipped to +-2 (pert 2) or uniform on (-1, 1) (pert 3, the peak
      // displacement is then at most amp h), keyed by the face's axis and
      // lattice position so both elements of an interior face draw the same
      // xi_f
      const int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
      const long fq[3] = {lat[axis] / q, lat[a0] / q, lat[a1] / q};
      double xi = normal_keyed(gp->seed, (uint64_t)axis + 1,
                               ((uint64_t)fq[0] << 21) ^ (uint64_t)fq[1], (uint64_t)fq[2]);
      xi = gp->pert == 3 ? std::erf(xi / std::sqrt(2.0)) : std::max(-2.0, std::min(2.0, xi));
      const double u = -1.0 + 2.0 * (double)(lat[a0] % q) / q;
      const double v = -1.0 + 2.0 * (double)(lat[a1] % q) / q;
      return gp->amp * h * xi * (1.0 - u * u) * (1.0 - v * v);
    }
    return 0.0;
  };

  parallel_chunks(N * N * N, [&](int e0, int e1) {
  std::vector<double> nodal(3 * n3), coef(n3);
  for (int ee = e0; ee < e1; ++ee) {
        const int i = ee % N, j = (ee / N) % N, k = ee / (N * N);
        const size_t e = (size_t)ee;
        double V[8][3];
        const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (int v = 0; v < 8; ++v) {
          V[v][0] = (i + off[v][0]) * h;
          V[v][1] = (j + off[v][1]) * h;
          V[v][2] = (k + off[v][2]) * h;
          if (vertices)
            for (int c = 0; c < 3; ++c) vertices[e * 24 + v * 3 + c] = (float)V[v][c];
        }
        if (geom) {
          // trilinear nodal values on the GLL grid
          for (int kk = 0; kk < ng; ++kk)
            for (int jj = 0; jj < ng; ++jj)
              for (int ii = 0; ii < ng; ++ii) {
                const double r = gll[ii], s = gll[jj], t = gll[kk];
                const size_t id = ((size_t)kk * ng + jj) * ng + ii;
                nodal[0 * n3 + id] = (i + 0.5 * (r + 1)) * h;
                nodal[1 * n3 + id] = (j + 0.5 * (s + 1)) * h;
                nodal[2 * n3 + id] = (k + 0.5 * (t + 1)) * h;
              }
          // curved faces: evenly spaced face nodes displaced along the face
          // normal axis, resampled to GLL and blended linearly into the volume
          // (ResampleQuadFaceLegendre + blend_face,
          // xml_fld_geometry_builder.cpp:99-233, 350-377).
          for (int lf = 0; lf < 6; ++lf) {
            const int fix = (lf == 0 || lf == 5) ? 2 : ((lf == 1 || lf == 3) ? 1 : 0);
            const int fixside = (lf == 2 || lf == 3 || lf == 5) ? 1 : 0;
            const int ua = (lf == 2 || lf == 4) ? 1 : 0;
            const int va = (lf == 0 || lf == 5) ? 1 : 2;
            // displacement on the evenly spaced face grid (zero on the face's
            // boundary ring: edges stay straight)
            std::vector<double> de((size_t)ng * ng, 0.0);
            bool any = false;
            for (int b = 1; b < q; ++b)
              for (int a = 1; a < q; ++a) {
                long lat[3];
                double w[3];
                const int base[3] = {i, j, k};
                lat[fix] = (long)(base[fix] + fixside) * q;
                lat[ua] = (long)base[ua] * q + a;
                lat[va] = (long)base[va] * q + b;
                for (int c = 0; c < 3; ++c) w[c] = (double)lat[c] / ((double)N * q);
                de[(size_t)b * ng + a] = displacement(fix, lat, w);
                if (de[(size_t)b * ng + a] != 0.0) any = true;
              }
            if (!any) continue;
            // resample to GLL face nodes: D = L de L^T
            std::vector<double> tmp((size_t)ng * ng, 0.0), dg((size_t)ng * ng, 0.0);
            for (int b = 0; b < ng; ++b)
              for (int a2 = 0; a2 < ng; ++a2) {
                double s = 0;
                for (int a = 0; a < ng; ++a) s += L[(size_t)a2 * ng + a] * de[(size_t)b * ng + a];
                tmp[(size_t)b * ng + a2] = s;
              }
            for (int b2 = 0; b2 < ng; ++b2)
              for (int a2 = 0; a2 < ng; ++a2) {
                double s = 0;
                for (int b = 0; b < ng; ++b) s += L[(size_t)b2 * ng + b] * tmp[(size_t)b * ng + a2];
                dg[(size_t)b2 * ng + a2] = s;
              }
            // blend along the normal coordinate with weight (1 -+ w)/2
            for (int b = 0; b < ng; ++b)
              for (int a = 0; a < ng; ++a) {
                const double dv = dg[(size_t)b * ng + a];
                if (dv == 0.0) continue;
                for (int m = 0; m < ng; ++m) {
                  int idx3[3];
                  idx3[ua] = a;
                  idx3[va] = b;
                  idx3[fix] = m;
                  const double wn = fixside ? 0.5 * (1.0 + gll[m]) : 0.5 * (1.0 - gll[m]);
                  const size_t id = ((size_t)idx3[2] * ng + idx3[1]) * ng + idx3[0];
                  nodal[(size_t)fix * n3 + id] += dv * wn;
                }
              }
          }
          for (int c = 0; c < 3; ++c) {
            tensor_fit(ng, inv, &nodal[(size_t)c * n3], coef.data());
            for (size_t m = 0; m < n3; ++m) geom[(e * 3 + c) * n3 + m] = (float)coef[m];
          }
        }
        if (field) {
          for (int kk = 0; kk < nf; ++kk)
            for (int jj = 0; jj < nf; ++jj)
              for (int ii = 0; ii < nf; ++ii) {
                const size_t m = ((size_t)kk * nf + jj) * nf + ii;
                const double g = normal_keyed(gp->seed ^ 0xF1E1Dull, e, m, 7);
                const double den = 1.0 + ii + jj + kk;
                field[e * f3 + m] = (float)(g / (den * den));
              }
        }
        if (adjacency) {
          const int nb[6][3] = {{0, 0, -1}, {0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, 0, 1}};
          const int opp[6] = {5, 3, 4, 1, 2, 0};
          for (int f = 0; f < 6; ++f) {
            const int ni = i + nb[f][0], nj = j + nb[f][1], nk = k + nb[f][2];
            int ne = -1;
            if (ni >= 0 && nj >= 0 && nk >= 0 && ni < N && nj < N && nk < N)
              ne = (int)(((size_t)nk * N + nj) * N + ni);
            adjacency[e * 12 + f * 2 + 0] = ne;
            adjacency[e * 12 + f * 2 + 1] = ne >= 0 ? opp[f] : -1;
          }
        }
      }
  });
  (void)kFaceCorners;
  return FEMRT_OK;
}

// -------------------------------------------------- multi-frequency fields

// Field u(x) = sum_w a_w sin(k_w . x + phi_w) / sqrt(sum_w a_w^2 / 2) (unit
// RMS): wave w has frequency f_w = 2^(w mod 4) (1, 2, 4, 8 periods over the
// unit cube) times a seeded factor in [1, 1.5), a seeded direction, amplitude
// 1 / f_w and a seeded phase.  Per element it is sampled at the nf^3 GLL
// nodes of the element's geometry map and fitted to nf modes per direction
// (the tensor modal fit used for geometry), so neighbouring elements agree
// up to the fit error.
int gen_field_waves(int n_elem, int ng, const float* geom, int nf, int n_waves, uint64_t seed,
                    float* field)
{
  if (n_elem < 0 || ng < 2 || ng > 16 || nf < 2 || nf > 16 || n_waves < 1 || n_waves > 64 || !geom || !field)
    return FEMRT_EINVAL;
  std::vector<double> kv(3 * (size_t)n_waves), ph(n_waves), am(n_waves);
  double norm = 0.0;
  for (int w = 0; w < n_waves; ++w) {
    const double u1 = 0.5 + 0.5 * std::erf(normal_keyed(seed, 11, (uint64_t)w, 0) / std::sqrt(2.0));
    const double f = (double)(1 << (w % 4)) * (1.0 + 0.5 * u1);
    double dir[3], dn = 0.0;
    for (int c = 0; c < 3; ++c) {
      dir[c] = normal_keyed(seed, 12, (uint64_t)w, (uint64_t)c);
      dn += dir[c] * dir[c];
    }
    dn = std::sqrt(std::max(dn, 1e-30));
    for (int c = 0; c < 3; ++c) kv[3 * (size_t)w + c] = 2.0 * M_PI * f * dir[c] / dn;
    ph[w] = 2.0 * M_PI * (0.5 + 0.5 * std::erf(normal_keyed(seed, 13, (uint64_t)w, 0) / std::sqrt(2.0)));
    am[w] = 1.0 / f;
    norm += 0.5 * am[w] * am[w];
  }
  norm = 1.0 / std::sqrt(norm);
  const std::vector<double> gll = gll_nodes(nf);
  std::vector<double> inv;
  if (!modal_fit_1d_matrix(nf, gll, inv)) return FEMRT_EINVAL;
  // geometry basis at the field's GLL nodes: Bm[a * ng + p] = phi_p(gll_a)
  std::vector<double> Bm((size_t)nf * ng);
  for (int a = 0; a < nf; ++a)
    for (int q = 0; q < ng; ++q) Bm[(size_t)a * ng + q] = modA(q, gll[a]);
  const size_t g3 = (size_t)ng * ng * ng, f3 = (size_t)nf * nf * nf;
  parallel_chunks(n_elem, [&](int e0, int e1) {
    std::vector<double> t1((size_t)ng * ng * nf), t2((size_t)ng * nf * nf), X(3 * f3), nodal(f3), coef(f3);
    for (int e = e0; e < e1; ++e) {
      // x at the GLL grid, sum-factorised over r, then s, then t
      for (int c = 0; c < 3; ++c) {
        const float* gc = geom + ((size_t)e * 3 + c) * g3;
        for (int r = 0; r < ng; ++r)
          for (int q = 0; q < ng; ++q)
            for (int i = 0; i < nf; ++i) {
              double acc = 0.0;
              for (int p0 = 0; p0 < ng; ++p0) acc += Bm[(size_t)i * ng + p0] * gc[((size_t)r * ng + q) * ng + p0];
              t1[((size_t)r * ng + q) * nf + i] = acc;
            }
        for (int r = 0; r < ng; ++r)
          for (int j = 0; j < nf; ++j)
            for (int i = 0; i < nf; ++i) {
              double acc = 0.0;
              for (int q = 0; q < ng; ++q) acc += Bm[(size_t)j * ng + q] * t1[((size_t)r * ng + q) * nf + i];
              t2[((size_t)r * nf + j) * nf + i] = acc;
            }
        for (int k = 0; k < nf; ++k)
          for (int j = 0; j < nf; ++j)
            for (int i = 0; i < nf; ++i) {
              double acc = 0.0;
              for (int r = 0; r < ng; ++r) acc += Bm[(size_t)k * ng + r] * t2[((size_t)r * nf + j) * nf + i];
              X[(size_t)c * f3 + ((size_t)k * nf + j) * nf + i] = acc;
            }
      }
      for (size_t m = 0; m < f3; ++m) {
        double u = 0.0;
        for (int w = 0; w < n_waves; ++w)
          u += am[w] * std::sin(kv[3 * (size_t)w] * X[m] + kv[3 * (size_t)w + 1] * X[f3 + m] +
                                kv[3 * (size_t)w + 2] * X[2 * f3 + m] + ph[w]);
        nodal[m] = u * norm;
      }
      tensor_fit(nf, inv, nodal.data(), coef.data());
      for (size_t m = 0; m < f3; ++m) field[(size_t)e * f3 + m] = (float)coef[m];
    }
  });
  return FEMRT_OK;
}

// ---------------------------------------------------------- Bernstein boxes

// Modal -> Bernstein change of basis for n modes, restricted to the
// sub-interval [a, b] of [-1, 1] and reparametrised to [-1, 1]:
//   phi_q(a + (b-a)(x+1)/2) = sum_m T[q][m] B_m(x),
//   B_m(x) = C(n-1,m) ((1+x)/2)^m ((1-x)/2)^(n-1-m).
// Bernstein coefficients of a tensor polynomial bound it on its domain
// (convex hull property), so their min/max is a guaranteed box, and a face's
// patch is bounded by the boundary layer of the net.  Splitting [-1,1] into S
// pieces per direction (S^3 sub-nets) tightens the bound: the hull of a
// sign-alternating bubble overshoots by up to ~4.5x per direction on the whole
// interval, and the overshoot shrinks quadratically with the piece length.
static std::vector<double> modal_to_bernstein(int n, double a, double b)
{
  std::vector<double> xs(n), VB((size_t)n * n), VP((size_t)n * n), iVB;
  for (int p = 0; p < n; ++p) xs[p] = n == 1 ? 0.0 : -1.0 + 2.0 * p / (n - 1);
  for (int p = 0; p < n; ++p) {
    const double wa = 0.5 * (1 + xs[p]), wb = 0.5 * (1 - xs[p]);
    for (int m = 0; m < n; ++m) {
      double c = 1.0;
      for (int r = 1; r <= m; ++r) c = c * (n - 1 - m + r) / r;
      VB[(size_t)p * n + m] = c * std::pow(wa, m) * std::pow(wb, n - 1 - m);
    }
    const double y = a + (b - a) * 0.5 * (xs[p] + 1.0);
    for (int q = 0; q < n; ++q) VP[(size_t)p * n + q] = modA(q, y);
  }
  invert(VB, n, iVB);
  // VP = VB T^T  ->  T[q][m] = sum_p iVB[m][p] VP[p][q]
  std::vector<double> T((size_t)n * n);
  for (int q = 0; q < n; ++q)
    for (int m = 0; m < n; ++m) {
      double s = 0;
      for (int p = 0; p < n; ++p) s += iVB[(size_t)m * n + p] * VP[(size_t)p * n + q];
      T[(size_t)q * n + m] = s;
    }
  return T;
}

// Per element, bounds of the field u over the reference cube: the extreme
// coefficients of u in the tensor Bernstein bases of the 8 half-cubes (a
// polynomial lies in the convex hull of its Bernstein coefficients), widened
// by 1e-5 of sum |c| for the FP32 evaluation error of u, rounded outward.
// out: lo, hi per element.
void field_range(int n_elem, int nf, const float* field, float* out)
{
  // on each of the 2 x 2 x 2 half-cubes (the hull over the whole cube is
  // loose for the oscillating interior modes: it excluded no isovalue on the
  // benchmark's P = 6 fields where the split hull excludes +-0.1 for ~6% of
  // the elements)
  const int n = nf, S = 2;
  std::vector<std::vector<double>> T(S);  // T[h][q][m]: modal q -> Bernstein m on half h
  for (int h = 0; h < S; ++h) T[h] = modal_to_bernstein(n, -1.0 + 2.0 * h / S, -1.0 + 2.0 * (h + 1) / S);
  const size_t n3 = (size_t)n * n * n;
  parallel_chunks(n_elem, [&](int e0, int e1) {
  std::vector<double> c(n3), t1(n3), t2(n3);
  for (int e = e0; e < e1; ++e) {
    double asum = 0.0;
    for (size_t m = 0; m < n3; ++m) {
      c[m] = field[(size_t)e * n3 + m];
      asum += std::fabs(c[m]);
    }
    double lo = 1e300, hi = -1e300;
    for (int h1 = 0; h1 < S; ++h1) {
      const double* A = T[h1].data();
      // along i (r), then j (s), then k (t); index (k*n + j)*n + i
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int mi = 0; mi < n; ++mi) {
            double acc = 0.0;
            for (int i = 0; i < n; ++i) acc += A[(size_t)i * n + mi] * c[((size_t)k * n + j) * n + i];
            t1[((size_t)k * n + j) * n + mi] = acc;
          }
      for (int h2 = 0; h2 < S; ++h2) {
        const double* B = T[h2].data();
        for (int k = 0; k < n; ++k)
          for (int mj = 0; mj < n; ++mj)
            for (int mi = 0; mi < n; ++mi) {
              double acc = 0.0;
              for (int j = 0; j < n; ++j) acc += B[(size_t)j * n + mj] * t1[((size_t)k * n + j) * n + mi];
              t2[((size_t)k * n + mj) * n + mi] = acc;
            }
        for (int h3 = 0; h3 < S; ++h3) {
          const double* Cm = T[h3].data();
          for (int mk = 0; mk < n; ++mk)
            for (int mj = 0; mj < n; ++mj)
              for (int mi = 0; mi < n; ++mi) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += Cm[(size_t)k * n + mk] * t2[((size_t)k * n + mj) * n + mi];
                lo = std::min(lo, acc);
                hi = std::max(hi, acc);
              }
        }
      }
    }
    const double pad = 1e-5 * asum + 1e-30;
    out[2 * (size_t)e] = std::nextafter((float)(lo - pad), -INFINITY);
    out[2 * (size_t)e + 1] = std::nextafter((float)(hi + pad), INFINITY);
  }
  });
}

// y = x transformed along one axis by A (A[q][m], q < d[axis], m < mo);
// index i + d0 (j + d1 k), i along r.  d is updated to y's dimensions.
static void along(const std::vector<double>& x, int* d, int axis, const double* A, int mo,
                  std::vector<double>& y)
{
  int od[3] = {d[0], d[1], d[2]};
  od[axis] = mo;
  y.assign((size_t)od[0] * od[1] * od[2], 0.0);
  for (int k = 0; k < od[2]; ++k)
    for (int j = 0; j < od[1]; ++j)
      for (int i = 0; i < od[0]; ++i) {
        const int o[3] = {i, j, k};
        double acc = 0.0;
        for (int q = 0; q < d[axis]; ++q) {
          int src[3] = {i, j, k};
          src[axis] = q;
          acc += A[(size_t)q * mo + o[axis]] * x[src[0] + (size_t)d[0] * (src[1] + (size_t)d[1] * src[2])];
        }
        y[i + (size_t)od[0] * (j + (size_t)od[1] * k)] = acc;
      }
  d[axis] = mo;
}

// Bernstein forms of the Modified_A basis and of its derivative on the S
// equal pieces of [-1, 1]: T[h][q][m] (stride n) and D[h][q][m] (stride
// n - 1).  On [a, b] the derivative of a degree-N Bernstein form has the
// degree-(N-1) coefficients N (beta_{m+1} - beta_m) / (b - a).
static void piece_bernstein(int n, int S, std::vector<std::vector<double>>& T,
                            std::vector<std::vector<double>>& D)
{
  const int N = n - 1;
  T.assign(S, {});
  D.assign(S, {});
  for (int h = 0; h < S; ++h) {
    const double a = -1.0 + 2.0 * h / S, b = -1.0 + 2.0 * (h + 1) / S;
    T[h] = modal_to_bernstein(n, a, b);
    D[h].assign((size_t)n * std::max(N, 1), 0.0);
    for (int q = 0; q < n; ++q)
      for (int m = 0; m < N; ++m)
        D[h][(size_t)q * N + m] = N * (T[h][(size_t)q * n + m + 1] - T[h][(size_t)q * n + m]) / (b - a);
  }
}

// Per element, K >= |J^-1|_inf (max row sum) everywhere in the reference
// cube, J = dx/drst, or 0 when no bound is found.  On each of the 4 x 4 x 4
// sub-cubes the 9 entries of J lie in intervals [J0 - R, J0 + R] (the
// extreme Bernstein coefficients of the geometry's partial derivatives).
// With M = J0^-1 and beta = |(|M| R)|_inf < 0.9, every J of the interval
// matrix has |J^-1|_inf <= |M|_inf / (1 - beta) (J^-1 = (I + M E)^-1 M,
// |M E| <= beta, a Neumann series).  K is the largest bound over the
// sub-cubes, plus 5%.  (Face bubbles shear J by up to ~60% of its diagonal
// across an element: on half-cubes beta came near 1 and K ~5x the true
// maximum; on quarter-cubes K is ~1.7x it on the 32^3 benchmark mesh.)
// Along a ray segment inside the element, |rst(t) - rst(t')|_inf <=
// K |d|_inf |t - t'|.
void geom_jinv_bound(int n_elem, int ng, const float* geom, float* out)
{
  const int n = ng, N = ng - 1, S = 4;
  std::vector<std::vector<double>> T, D;
  piece_bernstein(n, S, T, D);
  const size_t n3 = (size_t)n * n * n;
  const int S3 = S * S * S;
  parallel_chunks(n_elem, [&](int e0, int e1) {
    std::vector<double> c(n3), y1, y2, y3;
    std::vector<double> jlo((size_t)S3 * 9), jhi((size_t)S3 * 9);
    for (int e = e0; e < e1; ++e) {
      // intervals of J[a][ax] = dx_a/dr_ax on every sub-cube, by a tree of
      // transforms: along r per piece, then s, then t
      bool ok = N > 0;
      for (int a = 0; a < 3 && ok; ++a) {
        for (size_t m = 0; m < n3; ++m) c[m] = geom[((size_t)e * 3 + a) * n3 + m];
        for (int ax = 0; ax < 3; ++ax)
          for (int h0 = 0; h0 < S; ++h0) {
            int d1[3] = {n, n, n};
            along(c, d1, 0, ax == 0 ? D[h0].data() : T[h0].data(), ax == 0 ? N : n, y1);
            for (int h1 = 0; h1 < S; ++h1) {
              int d2[3] = {d1[0], d1[1], d1[2]};
              along(y1, d2, 1, ax == 1 ? D[h1].data() : T[h1].data(), ax == 1 ? N : n, y2);
              for (int h2 = 0; h2 < S; ++h2) {
                int d3[3] = {d2[0], d2[1], d2[2]};
                along(y2, d3, 2, ax == 2 ? D[h2].data() : T[h2].data(), ax == 2 ? N : n, y3);
                double lo = 1e300, hi = -1e300;
                for (double v : y3) {
                  lo = std::min(lo, v);
                  hi = std::max(hi, v);
                }
                const size_t h = ((size_t)h2 * S + h1) * S + h0;
                jlo[h * 9 + a * 3 + ax] = lo;
                jhi[h * 9 + a * 3 + ax] = hi;
              }
            }
          }
      }
      double K = 0.0;
      for (int h = 0; h < S3 && ok; ++h) {
        double J0[3][3], R[3][3];
        for (int a = 0; a < 3; ++a)
          for (int ax = 0; ax < 3; ++ax) {
            const double lo = jlo[(size_t)h * 9 + a * 3 + ax], hi = jhi[(size_t)h * 9 + a * 3 + ax];
            J0[a][ax] = 0.5 * (lo + hi);
            R[a][ax] = 0.5 * (hi - lo) * (1.0 + 1e-9) + 1e-12 * std::fabs(J0[a][ax]);
          }
        const double c00 = J0[1][1] * J0[2][2] - J0[1][2] * J0[2][1];
        const double c01 = J0[1][2] * J0[2][0] - J0[1][0] * J0[2][2];
        const double c02 = J0[1][0] * J0[2][1] - J0[1][1] * J0[2][0];
        const double det = J0[0][0] * c00 + J0[0][1] * c01 + J0[0][2] * c02;
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
          ok = false;
          break;
        }
        double M[3][3];
        M[0][0] = c00 / det;
        M[1][0] = c01 / det;
        M[2][0] = c02 / det;
        M[0][1] = (J0[0][2] * J0[2][1] - J0[0][1] * J0[2][2]) / det;
        M[1][1] = (J0[0][0] * J0[2][2] - J0[0][2] * J0[2][0]) / det;
        M[2][1] = (J0[0][1] * J0[2][0] - J0[0][0] * J0[2][1]) / det;
        M[0][2] = (J0[0][1] * J0[1][2] - J0[0][2] * J0[1][1]) / det;
        M[1][2] = (J0[0][2] * J0[1][0] - J0[0][0] * J0[1][2]) / det;
        M[2][2] = (J0[0][0] * J0[1][1] - J0[0][1] * J0[1][0]) / det;
        double mnorm = 0.0, beta = 0.0;
        for (int a = 0; a < 3; ++a) {
          double rs = 0.0, bs = 0.0;
          for (int b = 0; b < 3; ++b) {
            rs += std::fabs(M[a][b]);
            for (int k = 0; k < 3; ++k) bs += std::fabs(M[a][k]) * R[k][b];
          }
          mnorm = std::max(mnorm, rs);
          beta = std::max(beta, bs);
        }
        if (!(beta < 0.9)) {
          ok = false;
          break;
        }
        K = std::max(K, mnorm / (1.0 - beta));
      }
      out[e] = ok ? (float)(K * 1.05) : 0.f;
    }
  });
}

// Per element, a bound L of |u(a) - u(b)| / |a - b|_inf over the reference
// cube.  On each of the 2 x 2 x 2 half-cubes, G_c bounds |du/dc| (c = r, s,
// t) by the largest |Bernstein coefficient| of that partial derivative on
// the half-cube (a polynomial lies in the convex hull of its Bernstein
// coefficients); L is the largest G_r + G_s + G_t over the half-cubes.  On
// an interval [a, b] the derivative of a degree-N Bernstein form has the
// degree-(N-1) coefficients N (beta_{m+1} - beta_m) / (b - a).  Splitting
// matters: the Modified_A interior modes have large alternating Bernstein
// coefficients on the whole interval (at P = 8 the unsplit bound is ~150x
// the true gradient, split in halves ~1.5x).  5% margin: iterates up to
// 1e-3 outside the cube, FP32 evaluation.
void field_lipschitz(int n_elem, int nf, const float* field, float* out)
{
  const int n = nf, N = nf - 1, S = 2;
  std::vector<std::vector<double>> T, D;
  piece_bernstein(n, S, T, D);
  const size_t n3 = (size_t)n * n * n;
  parallel_chunks(n_elem, [&](int e0, int e1) {
    std::vector<double> c(n3), y1, y2, y3;
    for (int e = e0; e < e1; ++e) {
      for (size_t m = 0; m < n3; ++m) c[m] = field[(size_t)e * n3 + m];
      double L = 0.0;
      for (int h = 0; h < S * S * S && N > 0; ++h) {
        const int hx[3] = {h % S, (h / S) % S, h / (S * S)};
        double sum = 0.0;
        for (int axis = 0; axis < 3; ++axis) {
          int d[3] = {n, n, n};
          along(c, d, 0, axis == 0 ? D[hx[0]].data() : T[hx[0]].data(), axis == 0 ? N : n, y1);
          along(y1, d, 1, axis == 1 ? D[hx[1]].data() : T[hx[1]].data(), axis == 1 ? N : n, y2);
          along(y2, d, 2, axis == 2 ? D[hx[2]].data() : T[hx[2]].data(), axis == 2 ? N : n, y3);
          double g = 0.0;
          for (double v : y3) g = std::max(g, std::fabs(v));
          sum += g;
        }
        L = std::max(L, sum);
      }
      out[e] = (float)(L * 1.05);
    }
  });
}

static const int kBoxSplit = 3;
static_assert(kBoxSplit == 3, "face sub-patch layout assumes 3x3 sub-patches per face");

void bernstein_boxes(int n_elem, int ng, const float* geom, float* out)
{
  const int n = ng, S = kBoxSplit;
  const size_t n3 = (size_t)n * n * n;
  std::vector<std::vector<double>> T(S);
  for (int s = 0; s < S; ++s) T[s] = modal_to_bernstein(n, -1.0 + 2.0 * s / S, -1.0 + 2.0 * (s + 1) / S);
  parallel_chunks(n_elem, [&](int e0, int e1) {
  std::vector<double> cc(3 * n3), t1(3 * n3), t2(3 * n3), bz(3 * n3);
  for (int e = e0; e < e1; ++e) {
    double lo[FEMRT_NBOX][3], hi[FEMRT_NBOX][3];
    for (int b = 0; b < FEMRT_NBOX; ++b)
      for (int c = 0; c < 3; ++c) {
        lo[b][c] = 1e300;
        hi[b][c] = -1e300;
      }
    for (size_t m = 0; m < 3 * n3; ++m) cc[m] = geom[(size_t)e * 3 * n3 + m];
    for (int s1 = 0; s1 < S; ++s1) {
      const double* A = T[s1].data();
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < n; ++k)
          for (int j = 0; j < n; ++j)
            for (int m = 0; m < n; ++m) {
              double acc = 0;
              for (int i = 0; i < n; ++i) acc += cc[c * n3 + ((size_t)k * n + j) * n + i] * A[(size_t)i * n + m];
              t1[c * n3 + ((size_t)k * n + j) * n + m] = acc;
            }
      for (int s2 = 0; s2 < S; ++s2) {
        const double* B = T[s2].data();
        for (int c = 0; c < 3; ++c)
          for (int k = 0; k < n; ++k)
            for (int m2 = 0; m2 < n; ++m2)
              for (int m = 0; m < n; ++m) {
                double acc = 0;
                for (int j = 0; j < n; ++j) acc += t1[c * n3 + ((size_t)k * n + j) * n + m] * B[(size_t)j * n + m2];
                t2[c * n3 + ((size_t)k * n + m2) * n + m] = acc;
              }
        for (int s3 = 0; s3 < S; ++s3) {
          const double* Cm = T[s3].data();
          for (int c = 0; c < 3; ++c)
            for (int m3 = 0; m3 < n; ++m3)
              for (int m2 = 0; m2 < n; ++m2)
                for (int m = 0; m < n; ++m) {
                  double acc = 0;
                  for (int k = 0; k < n; ++k) acc += t2[c * n3 + ((size_t)k * n + m2) * n + m] * Cm[(size_t)k * n + m3];
                  bz[c * n3 + ((size_t)m3 * n + m2) * n + m] = acc;
                }
          const int sub[3] = {s1, s2, s3};
          for (int m3 = 0; m3 < n; ++m3)
            for (int m2 = 0; m2 < n; ++m2)
              for (int m = 0; m < n; ++m) {
                const int idx[3] = {m, m2, m3};
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = bz[c * n3 + ((size_t)m3 * n + m2) * n + m];
                for (int c = 0; c < 3; ++c) {
                  lo[0][c] = std::min(lo[0][c], v[c]);
                  hi[0][c] = std::max(hi[0][c], v[c]);
                }
                for (int f = 0; f < 6; ++f) {
                  const int fix = (f == 0 || f == 5) ? 2 : ((f == 1 || f == 3) ? 1 : 0);
                  const int ua = (f == 2 || f == 4) ? 1 : 0, va = (f == 0 || f == 5) ? 1 : 2;
                  const bool plus = (f == 2 || f == 3 || f == 5);
                  if (sub[fix] != (plus ? S - 1 : 0) || idx[fix] != (plus ? n - 1 : 0)) continue;
                  // face box and the box of the face's sub-patch (sub[ua], sub[va])
                  const int sp = 7 + 9 * f + 3 * sub[va] + sub[ua];
                  for (int c = 0; c < 3; ++c) {
                    lo[1 + f][c] = std::min(lo[1 + f][c], v[c]);
                    hi[1 + f][c] = std::max(hi[1 + f][c], v[c]);
                    lo[sp][c] = std::min(lo[sp][c], v[c]);
                    hi[sp][c] = std::max(hi[sp][c], v[c]);
                  }
                }
              }
        }
      }
    }
    // Face roots are accepted up to |u|, |v| <= 1 + 1e-4, a band just outside
    // the face; every box also covers that band: 2e-4 of the element extent
    // bounds 1e-4 |dx/du| for elements whose faces are not stretched beyond
    // twice their extent.
    double eext = 0;
    for (int c = 0; c < 3; ++c) eext = std::max(eext, hi[0][c] - lo[0][c]);
    for (int b = 0; b < FEMRT_NBOX; ++b) {
      double ext = 0;
      for (int c = 0; c < 3; ++c) ext = std::max(ext, hi[b][c] - lo[b][c]);
      for (int c = 0; c < 3; ++c) {
        const double pad = 2e-4 * eext + 1e-6 * ext +
                           1e-6 * std::max(std::fabs(lo[b][c]), std::fabs(hi[b][c])) + 1e-9;
        out[((size_t)e * FEMRT_NBOX + b) * 6 + c] = std::nextafter((float)(lo[b][c] - pad), -INFINITY);
        out[((size_t)e * FEMRT_NBOX + b) * 6 + 3 + c] = std::nextafter((float)(hi[b][c] + pad), INFINITY);
      }
    }
  }
  });
}

// ------------------------------------------------------------ binned SAH BVH

// Binned SAH over element-box centroids, 16 bins per axis, leaves of <= 4
// elements, split only when the SAH cost beats the leaf cost (the policy of
// BVH.cu:35-177).  Nodes: lo.xyz, a, hi.xyz, b; internal: a,b = children;
// leaf: a = first slot in prims, b = -count.
namespace {
struct Builder {
  const float* boxes;
  std::vector<int> idx;
  std::vector<float> cen;  // 3 per element
  BvhHost* out;

  void box_of(int s, int e, float* lo, float* hi)
  {
    for (int c = 0; c < 3; ++c) {
      lo[c] = INFINITY;
      hi[c] = -INFINITY;
    }
    for (int i = s; i < e; ++i) {
      const float* b = boxes + (size_t)idx[i] * FEMRT_BOX_STRIDE;
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], b[c]);
        hi[c] = std::max(hi[c], b[3 + c]);
      }
    }
  }
  static float area(const float* lo, const float* hi)
  {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dx * dz + dy * dz);
  }
  int emit(const float* lo, const float* hi, int a, int b)
  {
    const int id = (int)(out->nodes.size() / 8);
    float rec[8];
    rec[0] = lo[0]; rec[1] = lo[1]; rec[2] = lo[2];
    std::memcpy(&rec[3], &a, 4);
    rec[4] = hi[0]; rec[5] = hi[1]; rec[6] = hi[2];
    std::memcpy(&rec[7], &b, 4);
    out->nodes.insert(out->nodes.end(), rec, rec + 8);
    return id;
  }
  void set_children(int node, int a, int b)
  {
    std::memcpy(&out->nodes[(size_t)node * 8 + 3], &a, 4);
    std::memcpy(&out->nodes[(size_t)node * 8 + 7], &b, 4);
  }
  int build(int s, int e)
  {
    float lo[3], hi[3];
    box_of(s, e, lo, hi);
    const int count = e - s;
    if (count <= 4) return emit(lo, hi, s, -count);
    // centroid bounds
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = s; i < e; ++i)
      for (int c = 0; c < 3; ++c) {
        clo[c] = std::min(clo[c], cen[(size_t)idx[i] * 3 + c]);
        chi[c] = std::max(chi[c], cen[(size_t)idx[i] * 3 + c]);
      }
    const int NB = 16;
    float best = INFINITY;
    int bax = -1, bsplit = -1;
    const float parea = area(lo, hi);
    for (int ax = 0; ax < 3; ++ax) {
      const float ext = chi[ax] - clo[ax];
      if (!(ext > 1e-12f)) continue;
      int cnt[NB] = {0};
      float blo[NB][3], bhi[NB][3];
      for (int b = 0; b < NB; ++b)
        for (int c = 0; c < 3; ++c) {
          blo[b][c] = INFINITY;
          bhi[b][c] = -INFINITY;
        }
      for (int i = s; i < e; ++i) {
        int b = (int)(NB * (cen[(size_t)idx[i] * 3 + ax] - clo[ax]) / ext);
        b = std::max(0, std::min(NB - 1, b));
        ++cnt[b];
        const float* bx = boxes + (size_t)idx[i] * FEMRT_BOX_STRIDE;
        for (int c = 0; c < 3; ++c) {
          blo[b][c] = std::min(blo[b][c], bx[c]);
          bhi[b][c] = std::max(bhi[b][c], bx[3 + c]);
        }
      }
      // sweep
      float rA[NB], rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      int rC[NB], rc = 0;
      for (int b = NB - 1; b > 0; --b) {
        rc += cnt[b];
        for (int c = 0; c < 3; ++c) {
          rlo[c] = std::min(rlo[c], blo[b][c]);
          rhi[c] = std::max(rhi[c], bhi[b][c]);
        }
        rA[b] = rc ? area(rlo, rhi) : 0.f;
        rC[b] = rc;
      }
      float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      int lc = 0;
      for (int b = 0; b < NB - 1; ++b) {
        lc += cnt[b];
        for (int c = 0; c < 3; ++c) {
          llo[c] = std::min(llo[c], blo[b][c]);
          lhi[c] = std::max(lhi[c], bhi[b][c]);
        }
        if (lc == 0 || rC[b + 1] == 0) continue;
        const float cost = (area(llo, lhi) * lc + rA[b + 1] * rC[b + 1]) / parea;
        if (cost < best) {
          best = cost;
          bax = ax;
          bsplit = b;
        }
      }
    }
    if (bax < 0 || best >= (float)count) return emit(lo, hi, s, -count);
    const float ext = chi[bax] - clo[bax];
    auto mid_it = std::partition(idx.begin() + s, idx.begin() + e, [&](int id) {
      int b = (int)(NB * (cen[(size_t)id * 3 + bax] - clo[bax]) / ext);
      b = std::max(0, std::min(NB - 1, b));
      return b <= bsplit;
    });
    int mid = (int)(mid_it - idx.begin());
    if (mid == s || mid == e) return emit(lo, hi, s, -count);
    const int node = emit(lo, hi, 0, 0);
    const int l = build(s, mid);
    const int r = build(mid, e);
    set_children(node, l, r);
    return node;
  }
};
}  // namespace

void build_bvh_sah(int n_elem, const float* boxes, BvhHost& out)
{
  Builder b;
  b.boxes = boxes;
  b.out = &out;
  out.nodes.clear();
  b.idx.resize(n_elem);
  std::iota(b.idx.begin(), b.idx.end(), 0);
  b.cen.resize((size_t)n_elem * 3);
  for (int e = 0; e < n_elem; ++e)
    for (int c = 0; c < 3; ++c)
      b.cen[(size_t)e * 3 + c] = 0.5f * (boxes[(size_t)e * FEMRT_BOX_STRIDE + c] + boxes[(size_t)e * FEMRT_BOX_STRIDE + 3 + c]);
  if (n_elem > 0) b.build(0, n_elem);
  out.prims = b.idx;
}

}  // namespace femrt
