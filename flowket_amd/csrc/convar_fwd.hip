// ConvNetAutoregressive2D forward: host API (packing maps, planning,
// schedule table, dispatch) and the kernels whose five activation lattices are
// resident in LDS.  The kernel template is in convar_fwd_kernel.h.
// HBM traffic: H*W int8 spins in, 8 bytes out per configuration.
#include <map>
#include <utility>
#include <mutex>

#include "convar_fwd_kernel.h"

namespace fk {

// ---------------------------------------------------------------------------
__global__ void pack_gather_kernel(const float* __restrict__ src, const int32_t* __restrict__ map, int64_t n,
                                   int prec, void* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = map[i];
  const float v = k >= 0 ? src[k] : 0.0f;
  if (prec == FK_PREC_BF16) reinterpret_cast<__bf16*>(dst)[i] = (__bf16)v;
  else if (prec == FK_PREC_FP16) reinterpret_cast<_Float16*>(dst)[i] = (_Float16)v;
  else reinterpret_cast<float*>(dst)[i] = v;
}

// Activation buffers of P + ZERO_PIX pixels (five in LDS, or three with the
// horizontal stack in global scratch) and the int8 spins; the per-site
// log-probabilities live in V1.
static size_t fwd_lds_bytes(int P, int pixb, bool hg, int T) {
  return (size_t)(P + ZERO_PIX) * pixb * (hg ? 3 : 5) + (size_t)((P + 3) & ~3) + 4 * (size_t)T;
}

// One launch shape of the forward: loop order and waves per workgroup.
struct FwdSchedule {
  int sched, nwave;
};

// Schedules compiled for each (precision, channels); index 0 is the default.
static int fwd_schedules(int C, int prec, const FwdSchedule** out) {
  static const FwdSchedule tile_capable[] = {{0, 8}, {1, 4}, {1, 8}, {0, 4}};
  static const FwdSchedule taps_only[] = {{0, 8}, {0, 4}};
  if ((prec == FK_PREC_BF16 && C == 32) || (prec == FK_PREC_FP32 && C == 16)) {
    *out = tile_capable;
    return 4;
  }
  *out = taps_only;
  return 2;
}

// Selected schedule per network: lattice, depth, kernel channel width,
// activation and precision (two networks that differ in any of them keep
// their own choice, so tuning one never changes the other's summation order).
static std::mutex g_sched_mu;
static std::map<std::pair<uint64_t, uint64_t>, int> g_sched;

static std::pair<uint64_t, uint64_t> sched_key(int H, int W, int depth, int C, int act, int prec) {
  return {((uint64_t)(uint32_t)H << 32) | (uint64_t)(uint32_t)W,
          ((uint64_t)(uint32_t)depth << 32) | ((uint64_t)(uint32_t)C << 8) | ((uint64_t)(act & 15) << 4) |
              (uint64_t)(prec & 15)};
}

static int current_schedule(int H, int W, int depth, int C, int act, int prec) {
  std::lock_guard<std::mutex> g(g_sched_mu);
  auto it = g_sched.find(sched_key(H, W, depth, C, act, prec));
  return it == g_sched.end() ? 0 : it->second;
}

// Largest tile of configurations whose activation lattices fit in the 160 KB
// LDS with at most 4 (8 waves) or 8 (4 waves) 16-pixel tiles per wave.
// HWL: LDS pixels per configuration, HWC: computed pixels per configuration
// (they differ for row-reuse launches, whose halo rows are not computed).
static int plan_forward(int HWL, int HWC, int C, int prec, int NW, int64_t m, FwdPlan* plan, bool hg = false) {
  const int pixb = C * (prec == FK_PREC_FP32 ? 4 : 2);
  const size_t cap = FWD_DYN_LDS;
  const int maxt_cap = (NW == 4) ? 8 : 4;
  int best = 0;
  for (int T = 1; T <= 256; ++T) {
    const int npix = ((T * HWC + 15) / 16) * 16;
    if (fwd_lds_bytes(T * HWL, pixb, hg, T) > cap) break;
    const int per_wave = (npix / 16 + NW - 1) / NW;
    if (per_wave > maxt_cap) break;
    best = T;
    if ((int64_t)T >= m) break;
  }
  if (best == 0) return FK_ERR_UNSUPPORTED;
  plan->T = best;
  plan->npix = ((best * HWC + 15) / 16) * 16;
  plan->lds = fwd_lds_bytes(best * HWL, pixb, hg, best);
  plan->nwave = NW;
  const int per_wave = (plan->npix / 16 + NW - 1) / NW;
  plan->maxt = per_wave <= 2 ? 2 : (per_wave <= 4 ? 4 : 8);
  return FK_OK;
}

bool convar_supported(int channels, int precision) { return kernel_channels(channels, precision) > 0; }

}  // namespace fk

using namespace fk;

extern "C" {

int fk_convar_packed_sizes(int32_t depth, int32_t channels, int32_t precision, int64_t* n_weights, int64_t* n_aux) {
  FK_REQUIRE(n_weights && n_aux, "fk_convar_packed_sizes: null pointer");
  FK_REQUIRE(depth >= 1, "fk_convar_packed_sizes: depth must be >= 1");
  if (!convar_supported(channels, precision)) {
    set_error("ConvNetAutoregressive2D kernels: channels=%d precision=%d not implemented", channels, precision);
    return FK_ERR_UNSUPPORTED;
  }
  const ConvarLayout lay(kernel_channels(channels, precision), precision);
  const int nb = num_causal_blocks(depth);
  *n_weights = lay.total_frags(nb) * 64 * lay.E;
  *n_aux = lay.aux_total(nb);
  return FK_OK;
}

int fk_convar_pack_maps(int32_t depth, int32_t channels, int32_t precision, int32_t* wmap, int32_t* amap) {
  int64_t nw = 0, na = 0;
  int rc = fk_convar_packed_sizes(depth, channels, precision, &nw, &na);
  if (rc != FK_OK) return rc;
  FK_REQUIRE(wmap && amap, "fk_convar_pack_maps: null pointer");
  const int C = channels;
  const int CP = kernel_channels(channels, precision);
  const ConvarLayout lay(CP, precision);
  const int nb = num_causal_blocks(depth);
  for (int64_t i = 0; i < nw; ++i) wmap[i] = -1;
  for (int64_t i = 0; i < na; ++i) amap[i] = -1;
  // canonical offsets
  int64_t src = 0;
  struct Conv {
    int64_t koff, boff;
    int kh, kw, cin, cout;
  };
  auto take = [&](int kh, int kw, int cin, int cout) {
    Conv c{src, src + (int64_t)kh * kw * cin * cout, kh, kw, cin, cout};
    src = c.boff + cout;
    return c;
  };
  // kernel input channel -> channel of the C-channel network (-1: padding);
  // `cat`: the input is the concatenation [x | y] whose y half starts at CP/2
  auto real_cin = [&](int pc, const Conv& cv, bool cat) {
    if (!cat) return pc < cv.cin ? pc : -1;
    if (pc < CP / 2) return pc < C / 2 ? pc : -1;
    return pc - CP / 2 < C / 2 ? C / 2 + (pc - CP / 2) : -1;
  };
  auto fill = [&](int64_t frag_base, const Conv& cv, int mto, bool cat) {
    const int ntap = cv.kh * cv.kw;
    for (int tap = 0; tap < ntap; ++tap)
      for (int mt = 0; mt < mto; ++mt)
        for (int ks = 0; ks < lay.KS; ++ks) {
          const int64_t frag = frag_base + ((int64_t)tap * mto + mt) * lay.KS + ks;
          for (int lane = 0; lane < 64; ++lane) {
            const int m = lane % 16, g = lane / 16;
            for (int e = 0; e < lay.E; ++e) {
              const int pc = precision != FK_PREC_FP32 ? slot_channel_bf16(32 * ks + 8 * g + e)
                                                       : slot_channel_f32(ks, g);
              const int cin = real_cin(pc, cv, cat);
              const int cout = 16 * mt + m;
              const int64_t idx = (frag * 64 + lane) * lay.E + e;
              if (cout < cv.cout && cin >= 0)
                wmap[idx] = (int32_t)(cv.koff + ((int64_t)tap * cv.cin + cin) * cv.cout + cout);
            }
          }
        }
  };
  auto bias_copy = [&](int64_t dst, const Conv& cv) {
    for (int c = 0; c < cv.cout; ++c) amap[dst + c] = (int32_t)(cv.boff + c);
  };
  for (int b = 0; b < nb; ++b) {
    const int cin = b == 0 ? 1 : C;
    Conv v = take(3, 3, cin, C), h = take(1, 3, cin, C), x = take(1, 1, C, C / 2), y = take(1, 1, C, C / 2),
         cc = take(3, 3, C, C);
    const int64_t ab = lay.aux_bias(b);
    bias_copy(ab, v);
    bias_copy(ab + CP, h);
    bias_copy(ab + 2 * CP, x);
    bias_copy(ab + 2 * CP + CP / 2, y);
    bias_copy(ab + 3 * CP, cc);
    if (b == 0) {
      // (3,3,1,C) and (1,3,1,C) kernels, [tap][channel] with stride CP
      for (int t = 0; t < 9; ++t)
        for (int c = 0; c < C; ++c) amap[lay.aux_kv0(nb) + t * CP + c] = (int32_t)(v.koff + t * C + c);
      for (int t = 0; t < 3; ++t)
        for (int c = 0; c < C; ++c) amap[lay.aux_kh0(nb) + t * CP + c] = (int32_t)(h.koff + t * C + c);
    } else {
      fill(lay.off_v(b), v, lay.MT, false);
      fill(lay.off_h(b), h, lay.MT, false);
    }
    fill(lay.off_x(b), x, lay.MTH, false);
    fill(lay.off_y(b), y, lay.MTH, false);
    fill(lay.off_c(b), cc, lay.MT, true);
  }
  Conv head = take(1, 1, C, 4);
  fill(lay.head_base(nb), head, 1, false);
  for (int c = 0; c < 4; ++c) amap[lay.aux_head_bias(nb) + c] = (int32_t)(head.boff + c);
  return FK_OK;
}

int fk_pack_gather(const float* params, const int32_t* map, int64_t n, int32_t precision, void* dst, void* stream) {
  FK_REQUIRE(params && map && dst, "fk_pack_gather: null pointer");
  FK_REQUIRE(precision == FK_PREC_FP32 || precision == FK_PREC_BF16 || precision == FK_PREC_FP16,
             "fk_pack_gather: bad precision");
  if (n <= 0) return FK_OK;
  hipLaunchKernelGGL(pack_gather_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, params, map, n,
                     precision, dst);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"

namespace fk {

// Waves per workgroup of the HG kernels (tap-outer only): the selected
// schedule's, or 8 for a tile-outer one.
static int hg_waves(const FwdSchedule& sc) { return sc.sched == 0 ? sc.nwave : 8; }

// Bytes of global scratch the HG kernels need for a lattice (0 when the five
// lattices fit in LDS under the schedule currently selected for it).
static int64_t forward_workspace_bytes(int H, int W, int depth, int CP, int act, int prec, FwdPlan* hg_plan) {
  const FwdSchedule* table = nullptr;
  const int nsched = fwd_schedules(CP, prec, &table);
  int si = current_schedule(H, W, depth, CP, act, prec);
  if (si < 0 || si >= nsched) si = 0;
  FwdPlan pl;
  if (plan_forward(H * W, H * W, CP, prec, table[si].nwave, 1 << 30, &pl) == FK_OK) return 0;
  if (plan_forward(H * W, H * W, CP, prec, hg_waves(table[si]), 1 << 30, &pl, true) != FK_OK) return -1;
  if (hg_plan) *hg_plan = pl;
  const int pixb = CP * (prec == FK_PREC_FP32 ? 4 : 2);
  // the whole-lattice launch and every row-reuse launch (rows R0-2 .. H-1)
  int64_t need = 0;
  for (int r0 = 0; r0 < H; ++r0) {
    if (r0 == 1) continue;
    const int hr = r0 == 0 ? H : H - r0 + 2, hc = r0 == 0 ? H : H - r0;
    FwdPlan q;
    if (plan_forward(hr * W, hc * W, CP, prec, hg_waves(table[si]), 1 << 30, &q, true) != FK_OK) return -1;
    const int per_cu = fwd_per_cu(q);
    const int64_t b = (int64_t)device_cu_count() * per_cu * 2 * (int64_t)(q.T * hr * W + ZERO_PIX) * pixb;
    if (b > need) need = b;
  }
  return need;
}

// Row reuse (see FwdArgs): R0 = 0 with rcache and cache_write is the sample
// pass; R0 >= 2 a restricted pass over rows R0-2 .. H-1.
struct RowReuse {
  bool rows = false;  // the sample pass or a restricted pass (kernel variant with ROWS)
  int R0 = 0;
  uint8_t* rcache = nullptr;
  int cache_write = 0;
  float* prefix = nullptr;
  const int64_t* out_index = nullptr;
  const int64_t* m_range = nullptr;
};

static int convar_forward(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W,
                          int32_t depth, int32_t channels, int32_t activation, const int8_t* spins,
                          const int32_t* pairs, const int64_t* m_dev, int64_t m_count, int32_t conn_kind,
                          const int32_t* conn_a, const int32_t* conn_b, float* logpsi, float* cond_logp,
                          void* workspace, int64_t workspace_bytes, void* stream, const RowReuse& rr = RowReuse()) {
  FK_REQUIRE(wpacked && aux && spins && logpsi, "fk_convar_forward: null pointer");
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1, "fk_convar_forward: bad shape");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_convar_forward: bad activation");
  FK_REQUIRE(!pairs || (conn_a && conn_b), "fk_convar_forward: pairs need a connection table");
  if (!convar_supported(channels, precision)) {
    set_error("ConvNetAutoregressive2D kernels: channels=%d precision=%d not implemented", channels, precision);
    return FK_ERR_UNSUPPORTED;
  }
  if (m_count <= 0) return FK_OK;
  const int CP = kernel_channels(channels, precision);
  const FwdSchedule* table = nullptr;
  const int nsched = fwd_schedules(CP, precision, &table);
  int si = current_schedule(H, W, depth, CP, activation, precision);
  if (si < 0 || si >= nsched) si = 0;
  const FwdSchedule sc = table[si];
  FwdArgs a;
  a.w = wpacked;
  a.aux = aux;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.nblocks = num_causal_blocks(depth);
  a.spins = spins;
  a.pairs = pairs;
  a.m_dev = m_dev;
  a.m_count = m_count;
  a.conn_kind = conn_kind;
  a.conn_a = conn_a;
  a.conn_b = conn_b;
  a.logpsi = logpsi;
  a.cond = cond_logp;
  a.hscr = nullptr;
  a.R0 = rr.R0;
  a.ROW0 = rr.R0 >= 2 ? 2 : 0;
  a.HR = rr.R0 >= 2 ? H - rr.R0 + 2 : H;
  a.HC = rr.R0 >= 2 ? H - rr.R0 : H;
  a.rcache = rr.rcache;
  a.rcache_stride = (int64_t)a.nblocks * 2 * H * W * CP * (precision == FK_PREC_FP32 ? 4 : 2);
  a.prefix = rr.prefix;
  a.cache_write = rr.cache_write;
  a.out_index = rr.out_index;
  a.m_range = rr.m_range;
  const int HWL = a.HR * W, HWC = a.HC * W;
  hipStream_t s = (hipStream_t)stream;
  FwdPlan pl;
  if (plan_forward(HWL, HWC, CP, precision, sc.nwave, m_count, &pl) != FK_OK) {
    // horizontal stack in global scratch
    FwdPlan hp;
    const int64_t need = forward_workspace_bytes(H, W, depth, CP, activation, precision, &hp);
    if (need < 0) {
      set_error("ConvNetAutoregressive2D kernels: lattice %dx%d with %d channels does not fit the LDS even with the "
                "horizontal stack in global scratch", H, W, channels);
      return FK_ERR_UNSUPPORTED;
    }
    if (!workspace || workspace_bytes < need) {
      set_error("ConvNetAutoregressive2D kernels: lattice %dx%d with %d channels needs a %lld-byte workspace "
                "(fk_convar_forward_workspace_bytes, fk_convar_forward_ws)", H, W, channels, (long long)need);
      return FK_ERR_ARG;
    }
    if (plan_forward(HWL, HWC, CP, precision, hg_waves(sc), m_count, &pl, true) != FK_OK) return FK_ERR_UNSUPPORTED;
    if (fwd_grid(a, pl) * 2 * (int64_t)(pl.T * HWL + ZERO_PIX) * (CP * (precision == FK_PREC_FP32 ? 4 : 2)) >
        workspace_bytes) {
      set_error("ConvNetAutoregressive2D kernels: workspace too small for this launch");
      return FK_ERR_ARG;
    }
    a.T = pl.T;
    a.npix = pl.npix;
    a.hscr = reinterpret_cast<uint8_t*>(workspace);
    if (precision == FK_PREC_BF16) return dispatch_fwd<FK_PREC_BF16, true>(a, pl, CP, 0, pl.nwave, activation, rr.rows, s);
    if (precision == FK_PREC_FP16) return dispatch_fwd<FK_PREC_FP16, true>(a, pl, CP, 0, pl.nwave, activation, rr.rows, s);
    return dispatch_fwd<FK_PREC_FP32, true>(a, pl, CP, 0, pl.nwave, activation, rr.rows, s);
  }
  a.T = pl.T;
  a.npix = pl.npix;
  if (precision == FK_PREC_BF16) return dispatch_fwd<FK_PREC_BF16, false>(a, pl, CP, sc.sched, sc.nwave, activation, rr.rows, s);
  if (precision == FK_PREC_FP16) return dispatch_fwd<FK_PREC_FP16, false>(a, pl, CP, sc.sched, sc.nwave, activation, rr.rows, s);
  return dispatch_fwd<FK_PREC_FP32, false>(a, pl, CP, sc.sched, sc.nwave, activation, rr.rows, s);
}

}  // namespace fk

extern "C" {

int fk_convar_forward(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W, int32_t depth,
                      int32_t channels, int32_t activation, const int8_t* spins, const int32_t* pairs,
                      const int64_t* m_dev, int64_t m_count, int32_t conn_kind, const int32_t* conn_a,
                      const int32_t* conn_b, float* logpsi, float* cond_logp, void* stream) {
  return convar_forward(wpacked, aux, precision, H, W, depth, channels, activation, spins, pairs, m_dev, m_count,
                        conn_kind, conn_a, conn_b, logpsi, cond_logp, nullptr, 0, stream);
}

int fk_convar_forward_ws(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W,
                         int32_t depth, int32_t channels, int32_t activation, const int8_t* spins,
                         const int32_t* pairs, const int64_t* m_dev, int64_t m_count, int32_t conn_kind,
                         const int32_t* conn_a, const int32_t* conn_b, float* logpsi, float* cond_logp,
                         void* workspace, int64_t workspace_bytes, void* stream) {
  return convar_forward(wpacked, aux, precision, H, W, depth, channels, activation, spins, pairs, m_dev, m_count,
                        conn_kind, conn_a, conn_b, logpsi, cond_logp, workspace, workspace_bytes, stream);
}

int64_t fk_convar_rcache_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t precision) {
  if (H < 1 || W < 1 || depth < 1 || !convar_supported(channels, precision)) return -1;
  const int CP = kernel_channels(channels, precision);
  return (int64_t)num_causal_blocks(depth) * 2 * H * W * CP * (precision == FK_PREC_FP32 ? 4 : 2);
}

int fk_convar_forward_rows(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W,
                           int32_t depth, int32_t channels, int32_t activation, const int8_t* spins,
                           const int32_t* pairs, const int64_t* m_range, int64_t m_count, int32_t conn_kind,
                           const int32_t* conn_a, const int32_t* conn_b, int32_t row0, void* rcache,
                           int32_t cache_write, float* prefix, const int64_t* out_index, float* logpsi,
                           void* workspace, int64_t workspace_bytes, void* stream) {
  FK_REQUIRE(H >= 3 && H <= 64, "fk_convar_forward_rows: 3 <= H <= 64");
  FK_REQUIRE(row0 == 0 || (row0 >= 2 && row0 < H), "fk_convar_forward_rows: row0 must be 0 or in [2, H)");
  FK_REQUIRE(!cache_write || (row0 == 0 && !pairs), "fk_convar_forward_rows: the sample pass takes rows of spins");
  FK_REQUIRE(row0 == 0 || (pairs && rcache && prefix), "fk_convar_forward_rows: a row pass needs pairs and the cache");
  FK_REQUIRE(!cache_write || (rcache && prefix), "fk_convar_forward_rows: the sample pass needs the cache");
  RowReuse rr;
  rr.rows = row0 > 0 || cache_write;
  rr.R0 = row0;
  rr.rcache = reinterpret_cast<uint8_t*>(rcache);
  rr.cache_write = cache_write;
  rr.prefix = prefix;
  rr.out_index = out_index;
  rr.m_range = m_range;
  return convar_forward(wpacked, aux, precision, H, W, depth, channels, activation, spins, pairs, nullptr, m_count,
                        conn_kind, conn_a, conn_b, logpsi, nullptr, workspace, workspace_bytes, stream, rr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward: the global-scratch workspace of the large-lattice
// kernels is a file-scope device array (one copy per GPU, held until the
// process ends).  Host code never passes the array's name to a kernel: the
// per-device address comes from hipGetSymbolAddress, taken once per device
// with that device current.
namespace {
constexpr int64_t NOALLOC_SCRATCH_BYTES = 32ll << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;

int scratch_address(void** out) {
  int dev = 0;
  FK_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return FK_OK;
  }
  void* ptr = nullptr;
  FK_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = ptr;
  *out = ptr;
  return FK_OK;
}
}  // namespace

extern "C" {

int64_t fk_convar_forward_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int fk_convar_forward_noalloc(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W,
                              int32_t depth, int32_t channels, int32_t activation, const int8_t* spins,
                              const int32_t* pairs, const int64_t* m_dev, int64_t m_count, int32_t conn_kind,
                              const int32_t* conn_a, const int32_t* conn_b, float* logpsi, float* cond_logp,
                              void* stream) {
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1, "fk_convar_forward_noalloc: bad shape");
  if (convar_supported(channels, precision)) {
    const int64_t need = forward_workspace_bytes(H, W, depth, kernel_channels(channels, precision), activation,
                                                 precision, nullptr);
    FK_REQUIRE(need <= NOALLOC_SCRATCH_BYTES,
               "fk_convar_forward_noalloc: a %dx%d lattice with %d channels needs %lld bytes of scratch, more "
               "than the %lld-byte static buffer (use fk_convar_forward_ws)",
               H, W, channels, (long long)need, (long long)NOALLOC_SCRATCH_BYTES);
  }
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != FK_OK) return rc;
  return convar_forward(wpacked, aux, precision, H, W, depth, channels, activation, spins, pairs, m_dev, m_count,
                        conn_kind, conn_a, conn_b, logpsi, cond_logp, ws, NOALLOC_SCRATCH_BYTES, stream);
}

int64_t fk_convar_forward_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t activation,
                                          int32_t precision) {
  if (H < 1 || W < 1 || depth < 1 || !convar_supported(channels, precision)) return -1;
  return forward_workspace_bytes(H, W, depth, kernel_channels(channels, precision), activation, precision, nullptr);
}

int32_t fk_convar_num_schedules(int32_t channels, int32_t precision) {
  if (!convar_supported(channels, precision)) return 0;
  const FwdSchedule* t = nullptr;
  return fwd_schedules(kernel_channels(channels, precision), precision, &t);
}

int fk_convar_schedule_info(int32_t channels, int32_t precision, int32_t schedule, int32_t* tile_outer,
                            int32_t* waves) {
  FK_REQUIRE(tile_outer && waves, "fk_convar_schedule_info: null pointer");
  const int n = fk_convar_num_schedules(channels, precision);
  FK_REQUIRE(schedule >= 0 && schedule < n, "fk_convar_schedule_info: schedule out of range");
  const FwdSchedule* t = nullptr;
  fwd_schedules(kernel_channels(channels, precision), precision, &t);
  *tile_outer = t[schedule].sched;
  *waves = t[schedule].nwave;
  return FK_OK;
}

int fk_convar_set_schedule(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t activation,
                           int32_t precision, int32_t schedule) {
  const int n = fk_convar_num_schedules(channels, precision);
  FK_REQUIRE(n > 0, "fk_convar_set_schedule: unsupported channels/precision");
  FK_REQUIRE(schedule >= -1 && schedule < n, "fk_convar_set_schedule: schedule out of range");
  const int CP = kernel_channels(channels, precision);
  std::lock_guard<std::mutex> g(g_sched_mu);
  if (schedule < 0) g_sched.erase(sched_key(H, W, depth, CP, activation, precision));
  else g_sched[sched_key(H, W, depth, CP, activation, precision)] = schedule;
  return FK_OK;
}

int32_t fk_convar_get_schedule(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t activation,
                               int32_t precision) {
  return current_schedule(H, W, depth, kernel_channels(channels, precision), activation, precision);
}

}  // extern "C"
