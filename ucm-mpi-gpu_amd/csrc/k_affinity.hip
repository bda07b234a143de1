// Stage (c): intervening-contour affinity W as a radius-5 stencil, its degree
// D and the normalised operator M = D^-1/2 W D^-1/2 used by the eigensolver.
//
// Reference: spectralPb.cpp:55-69 -> buildW (buildW.cpp:23-212) ->
// Group::computeSupport / interveningContour / ic_walk (buildW/src/ic.cc),
// computeAffinities2 (affinity.cc:58-156) and SMatrix::symmetrize
// (smatrix.cc:95-117).  The walks are replayed once on the host per border
// configuration (gpb::ic_program): each connection of a pixel to a disc
// neighbour becomes a short list of pixel offsets whose mPb maximum is the
// intervening contour.  The kernel stages an mPb tile with a 5-pixel halo in
// LDS and evaluates, for every upper disc offset, both directed affinities and
// their average; the degree kernel then forms D and D^-1/2.  W is stored once per unordered pair: Wup[k][p] for the 40
// offsets with dy > 0 or (dy == 0, dx > 0), in a zero-padded layout (5-pixel
// halo) so that every stencil access downstream is in bounds.
#include <mutex>

#include "gpb_pipeline.hpp"

namespace gpb {

constexpr int kAffTX = 32, kAffTY = 8;
constexpr int kAffTileW = kAffTX + 2 * kICRadius, kAffTileH = kAffTY + 2 * kICRadius;
constexpr int kNumConfigs = 6 * 6 * 6 * 6;

struct ICConnDev {
  uint32_t site_start;
  uint8_t site_count;
  uint8_t valid;   // neighbour inside the clipped window
  uint8_t found;   // support entry present
  uint8_t pad;
};

struct AffTables {
  DevBuf<ICConnDev> conn;    // [config][81]
  DevBuf<int16_t> site;      // LDS offsets relative to the pixel (dy * kAffTileW + dx)
  int updx[kStencilHalf], updy[kStencilHalf], upk[kStencilHalf];  // upper offsets and their disc index
};

namespace {
std::map<int, std::unique_ptr<AffTables>> g_aff;
std::mutex g_aff_mu;
}  // namespace

static AffTables& aff_tables(Context& c) {
  std::lock_guard<std::mutex> lk(g_aff_mu);
  auto& slot = g_aff[c.device];
  if (slot) return *slot;
  auto t = std::make_unique<AffTables>();
  int ddx[kStencilFull], ddy[kStencilFull];
  disc_offsets(ddx, ddy);
  std::vector<ICConnDev> conn(size_t(kNumConfigs) * kStencilFull);
  std::vector<int16_t> sites;
  for (int cfg = 0; cfg < kNumConfigs; cfg++) {
    const int l = cfg / 216, r = (cfg / 36) % 6, tp = (cfg / 6) % 6, b = cfg % 6;
    // Windows clipped to a single row or column (images 1 pixel wide or high)
    // have no valid walk set (ic.cc asserts); such images are rejected below.
    if ((l == 0 && r == 0) || (tp == 0 && b == 0)) continue;
    const ICProgram& prog = ic_program(l, r, tp, b);
    for (const ICConnection& cn : prog.conn) {
      int k = -1;
      for (int q = 0; q < kStencilFull; q++)
        if (ddx[q] == cn.dx && ddy[q] == cn.dy) k = q;
      if (k < 0) throw std::logic_error("ic program offset outside the disc");
      ICConnDev d{};
      d.site_start = uint32_t(sites.size());
      d.site_count = cn.site_count;
      d.valid = 1;
      d.found = cn.found;
      for (int s = 0; s < cn.site_count; s++)
        sites.push_back(int16_t(prog.site_dy[cn.site_start + s] * kAffTileW + prog.site_dx[cn.site_start + s]));
      conn[size_t(cfg) * kStencilFull + k] = d;
    }
  }
  t->conn.alloc(conn.size());
  t->site.alloc(std::max<size_t>(1, sites.size()));
  GPB_HIP(hipMemcpy(t->conn.p, conn.data(), conn.size() * sizeof(ICConnDev), hipMemcpyHostToDevice));
  if (!sites.empty()) GPB_HIP(hipMemcpy(t->site.p, sites.data(), sites.size() * sizeof(int16_t), hipMemcpyHostToDevice));
  upper_offsets(t->updx, t->updy);
  for (int j = 0; j < kStencilHalf; j++)
    for (int q = 0; q < kStencilFull; q++)
      if (ddx[q] == t->updx[j] && ddy[q] == t->updy[j]) t->upk[j] = q;
  slot = std::move(t);
  return *slot;
}

struct AffParams {
  const float* mpb;      // H x W (nonmax mPb in [0,1])
  const ICConnDev* conn;
  const int16_t* site;
  int updx[kStencilHalf], updy[kStencilHalf], upk[kStencilHalf];
  int W, H, Wp;          // Wp = W + 10 (padded row length)
  size_t Np;
  void* wup;             // [40][Np] symmetric affinities, zero halo (float, or double in the FP64 variant)
};

__device__ __forceinline__ int clip_code(int lo_dist) { return lo_dist < kICRadius ? lo_dist : kICRadius; }

__global__ __launch_bounds__(256) void to_f32_kernel(const double* __restrict__ in, size_t n, float* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    out[i] = float(in[i]);
}

// Affinity of a directed connection from the IC max (affinity.cc:40-52,
// 121-143): pss = C_IC_SS(1 - icsim), W = exp(-(1 - pss) / 0.1).
__device__ __forceinline__ float directed_w(float maxpb, bool found) {
  const float sim = 1.0f - maxpb;
  const float icsim = found ? sim : 0.0f;
  const float ic = 1.0f - icsim;
  float val = 1.8682f;
  val = float(double(val) + (double(ic) / 0.3130) * -1.3113);
  float post = float(1.0 / (1.0 + exp(-double(val))));
  if (!isfinite(post)) post = 0.f;
  if (post < 0.f) post = 0.f;
  if (post > 1.f) post = 1.f;
  const float e = -(1.0f - post) / 0.1f;
  return float(exp(double(e)));
}

// WT: storage of W.  The pair's two directed affinities are FP32 values (as
// affinity.cc:121-143 computes them) averaged in FP64 (smatrix.cc:95-117);
// float storage rounds that mean once more.
template <typename WT>
__global__ __launch_bounds__(kAffTX * kAffTY) void affinity_kernel(AffParams P) {
  __shared__ float tile[kAffTileH * kAffTileW];
  const int tid = threadIdx.y * kAffTX + threadIdx.x;
  const int x0 = blockIdx.x * kAffTX - kICRadius, y0 = blockIdx.y * kAffTY - kICRadius;
  for (int i = tid; i < kAffTileH * kAffTileW; i += kAffTX * kAffTY) {
    const int ty = i / kAffTileW, tx = i - ty * kAffTileW;
    const int gx = x0 + tx, gy = y0 + ty;
    tile[i] = (gx >= 0 && gx < P.W && gy >= 0 && gy < P.H) ? P.mpb[size_t(gy) * P.W + gx] : 0.f;
  }
  __syncthreads();
  const int x = blockIdx.x * kAffTX + threadIdx.x, y = blockIdx.y * kAffTY + threadIdx.y;
  if (x >= P.W || y >= P.H) return;
  const int cfg_p = ((clip_code(x) * 6 + clip_code(P.W - 1 - x)) * 6 + clip_code(y)) * 6 + clip_code(P.H - 1 - y);
  const int cp = (threadIdx.y + kICRadius) * kAffTileW + threadIdx.x + kICRadius;
  const size_t pp = size_t(y + kICRadius) * P.Wp + (x + kICRadius);
  for (int j = 0; j < kStencilHalf; j++) {
    const int dx = P.updx[j], dy = P.updy[j];
    const int qx = x + dx, qy = y + dy;
    double w = 0.0;
    if (qx >= 0 && qx < P.W && qy >= 0 && qy < P.H) {
      const ICConnDev a = P.conn[size_t(cfg_p) * kStencilFull + P.upk[j]];
      float m1 = 0.f;
      for (int s = 0; s < a.site_count; s++) m1 = fmaxf(m1, tile[cp + P.site[a.site_start + s]]);
      const int cfg_q = ((clip_code(qx) * 6 + clip_code(P.W - 1 - qx)) * 6 + clip_code(qy)) * 6 + clip_code(P.H - 1 - qy);
      const ICConnDev b = P.conn[size_t(cfg_q) * kStencilFull + (kStencilFull - 1 - P.upk[j])];
      const int cq = cp + dy * kAffTileW + dx;
      float m2 = 0.f;
      for (int s = 0; s < b.site_count; s++) m2 = fmaxf(m2, tile[cq + P.site[b.site_start + s]]);
      const float w1 = directed_w(m1, a.found), w2 = directed_w(m2, b.found);
      w = 0.5 * (double(w1) + double(w2));
    }
    static_cast<WT*>(P.wup)[size_t(j) * P.Np + pp] = WT(w);
  }
}

// D_p = 1 + sum of the 80 neighbour affinities; s_p = D_p^-1/2; diag = s_p^2.
template <typename WT>
__global__ __launch_bounds__(256) void degree_kernel(const WT* __restrict__ wup, int W, int H, int Wp, size_t Np,
                                                     const int* __restrict__ offs /* 40 padded offsets */,
                                                     WT* __restrict__ D, WT* __restrict__ s) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const size_t pp = size_t(y + kICRadius) * Wp + (x + kICRadius);
  double acc = 1.0;
  for (int j = 0; j < kStencilHalf; j++) {
    acc += double(wup[size_t(j) * Np + pp]);
    acc += double(wup[size_t(j) * Np + pp - offs[j]]);
  }
  D[pp] = WT(acc);
  s[pp] = WT(1.0 / sqrt(acc));
}

void build_stencil(Context& c, const float* d_mpb, int W, int H, Stencil& st, const std::string& tag) {
  if (W < 2 || H < 2) throw ArgError("the affinity needs an image at least 2 pixels wide and high");
  AffTables& t = aff_tables(c);
  st.W = W;
  st.H = H;
  st.Wp = W + 2 * kICRadius;
  st.Hp = H + 2 * kICRadius;
  st.Np = size_t(st.Wp) * st.Hp;
  if (tag.empty()) c.last_run_W = 0;  // the untagged stencil of the last whole-path run is replaced
  st.wup = c.ws<float>(tag + "st.wup", kStencilHalf * st.Np);
  st.D = c.ws<float>(tag + "st.D", st.Np);
  st.s = c.ws<float>(tag + "st.s", st.Np);
  st.offs = c.ws<int>(tag + "st.offs", kStencilHalf);
  int offs[kStencilHalf];
  for (int j = 0; j < kStencilHalf; j++) offs[j] = t.updy[j] * st.Wp + t.updx[j];
  GPB_HIP(hipMemcpyAsync(st.offs, offs, sizeof(offs), hipMemcpyHostToDevice, c.stream));
  for (int j = 0; j < kStencilHalf; j++) st.off_host[j] = offs[j];
  // zero halos: the kernels below write interior pixels only
  launch_fill(c, st.wup, kStencilHalf * st.Np, 0.f);
  launch_fill(c, st.D, st.Np, 0.f);
  launch_fill(c, st.s, st.Np, 0.f);
  AffParams P{};
  P.mpb = d_mpb;
  P.conn = t.conn.p;
  P.site = t.site.p;
  for (int j = 0; j < kStencilHalf; j++) {
    P.updx[j] = t.updx[j];
    P.updy[j] = t.updy[j];
    P.upk[j] = t.upk[j];
  }
  P.W = W;
  P.H = H;
  P.Wp = st.Wp;
  P.Np = st.Np;
  P.wup = st.wup;
  KScope ks(c, kKAffinity, double(W) * H * (4 + 4 * kStencilHalf));
  hipLaunchKernelGGL(affinity_kernel<float>, dim3(div_up(W, kAffTX), div_up(H, kAffTY)), dim3(kAffTX, kAffTY), 0,
                     c.stream, P);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(degree_kernel<float>, dim3(div_up(W, 256), H), dim3(256), 0, c.stream, st.wup, W, H, st.Wp, st.Np,
                     st.offs, st.D, st.s);
  GPB_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void fill_d_kernel(double* __restrict__ p, size_t n, double v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

void launch_fill_f64(Context& c, double* p, size_t n, double v) {
  if (!n) return;
  hipLaunchKernelGGL(fill_d_kernel, dim3(std::min<unsigned>(4096, div_up(n, 256))), dim3(256), 0, c.stream, p, n, v);
  GPB_LAUNCH_CHECK();
}

// The same stencil with W, D and D^-1/2 stored in FP64, from an FP64 mPb
// rounded to FP32 for the walks (the lattice of buildW.cpp:57-117 is FP32).
void build_stencil_f64(Context& c, const double* d_mpb, int W, int H, Stencil64& st) {
  if (W < 2 || H < 2) throw ArgError("the affinity needs an image at least 2 pixels wide and high");
  AffTables& t = aff_tables(c);
  st.W = W;
  st.H = H;
  st.Wp = W + 2 * kICRadius;
  st.Hp = H + 2 * kICRadius;
  st.Np = size_t(st.Wp) * st.Hp;
  st.wup = c.ws<double>("st64.wup", kStencilHalf * st.Np);
  st.D = c.ws<double>("st64.D", st.Np);
  st.s = c.ws<double>("st64.s", st.Np);
  st.offs = c.ws<int>("st64.offs", kStencilHalf);
  float* m32 = c.ws<float>("st64.mpb32", size_t(W) * H);
  hipLaunchKernelGGL(to_f32_kernel, dim3(std::min<unsigned>(4096, div_up(size_t(W) * H, 256))), dim3(256), 0, c.stream,
                     d_mpb, size_t(W) * H, m32);
  GPB_LAUNCH_CHECK();
  int offs[kStencilHalf];
  for (int j = 0; j < kStencilHalf; j++) offs[j] = t.updy[j] * st.Wp + t.updx[j];
  GPB_HIP(hipMemcpyAsync(st.offs, offs, sizeof(offs), hipMemcpyHostToDevice, c.stream));
  for (int j = 0; j < kStencilHalf; j++) st.off_host[j] = offs[j];
  launch_fill_f64(c, st.wup, kStencilHalf * st.Np, 0.0);
  launch_fill_f64(c, st.D, st.Np, 0.0);
  launch_fill_f64(c, st.s, st.Np, 0.0);
  AffParams P{};
  P.mpb = m32;
  P.conn = t.conn.p;
  P.site = t.site.p;
  for (int j = 0; j < kStencilHalf; j++) {
    P.updx[j] = t.updx[j];
    P.updy[j] = t.updy[j];
    P.upk[j] = t.upk[j];
  }
  P.W = W;
  P.H = H;
  P.Wp = st.Wp;
  P.Np = st.Np;
  P.wup = st.wup;
  KScope ks(c, kKAffinity, double(W) * H * (8 + 8 * kStencilHalf));
  hipLaunchKernelGGL(affinity_kernel<double>, dim3(div_up(W, kAffTX), div_up(H, kAffTY)), dim3(kAffTX, kAffTY), 0,
                     c.stream, P);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(degree_kernel<double>, dim3(div_up(W, 256), H), dim3(256), 0, c.stream, st.wup, W, H, st.Wp, st.Np,
                     st.offs, st.D, st.s);
  GPB_LAUNCH_CHECK();
}

void stencil_upper_offsets(int* dx, int* dy) { upper_offsets(dx, dy); }

}  // namespace gpb
