// Host side of the BLDFM C ABI: plan construction (validation, buffer
// layout, tile tables, FFT factorisation and twiddles), execution, profiles
// and the plain 2-D FFT.
#include <math.h>
#include <string.h>

#include <cmath>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bldfm_amd.h"
#include "bldfm_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(BLDFM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Restores the caller's current device on scope exit.
struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

constexpr int kMaxE = BLDFM_FFT_MAXE;
constexpr int kGridCap = 1 << 20;

int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// ------------------------------------------------------------- FFT planning
// Radix order: 8s, then one 4 or 2, then 3s, 5s, then remaining primes.
bool factorize(int n, FftPlan* P) {
  memset(P, 0, sizeof(*P));
  P->n = n;
  P->nstage = 0;
  int m = n;
  auto push = [&](int r) {
    if (P->nstage >= BLDFM_MAX_RADIX_STAGES) return false;
    P->radix[P->nstage++] = r;
    return true;
  };
  while (m % 8 == 0) { if (!push(8)) return false; m /= 8; }
  if (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
  if (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) { if (!push(p)) return false; m /= p; }
    if (p * p > m && m > 1) { if (!push(m)) return false; m = 1; }
  }
  int64_t ns = 1;
  for (int s = 0; s < P->nstage; ++s) {
    P->stride[s] = (int32_t)(n / (ns * P->radix[s]));
    P->magic[s] = ns >= 2 ? (uint32_t)((1ull << 32) / (uint64_t)ns + 1ull) : 0u;
    ns *= P->radix[s];
  }
  return true;
}

// W_n^t = e^{-2 pi i t/n} in extended precision, with exact values on the
// axes and exact conjugate symmetry W^{n-t} = conj W^t.
void twiddles(int n, double2* w) {
  const long double two_pi = 2.0L * acosl(-1.0L);
  for (int t = 0; t <= n / 2; ++t) {
    double c, s;
    if ((4LL * t) % n == 0) {
      const int q = (int)((4LL * t) / n);  // quarter turns
      const double cs[4] = {1.0, 0.0, -1.0, 0.0};
      const double sn[4] = {0.0, -1.0, 0.0, 1.0};
      c = cs[q & 3];
      s = sn[q & 3];
    } else {
      const long double a = two_pi * (long double)t / (long double)n;
      c = (double)cosl(a);
      s = (double)(-sinl(a));
    }
    w[t].x = c;
    w[t].y = s;
  }
  for (int t = n / 2 + 1; t < n; ++t) {
    w[t].x = w[n - t].x;
    w[t].y = -w[n - t].y;
  }
}

int fft_threads_for(int n) {
  const int e = n <= BLDFM_FFT_SMALL_N ? 8 : kMaxE;
  int t = (n + e - 1) / e;
  t = (int)align_up(std::max(t, 64), 64);
  return t;
}

int max_lds_bytes(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
    v = 64 * 1024;
  return v;
}

// Longest transform of one workgroup: buffer and twiddle tables in LDS.
int max_fft_len(int device) {
  const int lds = max_lds_bytes(device);
  int n = std::min(lds / 16, 1024 * kMaxE);
  while (n > 1 && fft_lds_bytes(n) > lds) --n;
  return n;
}

// Lengths above one workgroup's LDS limit run as four-step transforms of two
// factors that fit; a length without such a factor pair (a prime above the
// limit, say) is refused.  Returns the split factor n1 (0: no split) or -1.
int check_fft_len(int n, int device, const char* what, int* split = nullptr) {
  const int limit = max_fft_len(device);
  const int s = fft_split(n, limit);
  if (s < 0)
    return fail(BLDFM_ELIMIT, std::string(what) + " = " + std::to_string(n) +
                                  " exceeds the single-workgroup FFT limit (" + std::to_string(limit) +
                                  " points) and has no split into two factors within it");
  if (split) *split = s;
  return 0;
}

// The length whose LDS transform a plan of length n runs: n, or its larger factor.
int lds_len(int n, int split) { return split > 0 ? std::max(split, n / split) : n; }

// Plans of the plain 2-D FFT, per (device, n): one device block holding the
// FftPlan followed by its n twiddles.  Least recently used blocks are freed
// once a device's blocks exceed kTwCacheBytes; the device is synchronised
// first, so no kernel still in flight can read a freed block.
struct TwBlock {
  char* ptr;
  size_t bytes;
  uint64_t last_use;
};
std::mutex g_tw_mu;
std::map<std::pair<int, int>, TwBlock> g_tw_cache;
uint64_t g_tw_clock = 0;
constexpr size_t kPlanHdr = 256;                // FftPlan padded to 256 bytes
constexpr size_t kTwCacheBytes = 16u << 20;     // per device
static_assert(sizeof(FftPlan) <= kPlanHdr, "FftPlan must fit its header");

size_t tw_cache_bytes(int device) {
  size_t b = 0;
  for (const auto& kv : g_tw_cache)
    if (kv.first.first == device) b += kv.second.bytes;
  return b;
}

// Frees least recently used blocks of `device` (except `keep`) until `need`
// more bytes fit under the cap.  Called with g_tw_mu held and the device current.
void tw_evict(int device, size_t need, std::pair<int, int> keep) {
  size_t used = tw_cache_bytes(device);
  bool synced = false;
  while (used + need > kTwCacheBytes) {
    auto victim = g_tw_cache.end();
    for (auto it = g_tw_cache.begin(); it != g_tw_cache.end(); ++it)
      if (it->first.first == device && it->first != keep &&
          (victim == g_tw_cache.end() || it->second.last_use < victim->second.last_use))
        victim = it;
    if (victim == g_tw_cache.end()) break;
    if (!synced) {
      (void)hipDeviceSynchronize();
      synced = true;
    }
    (void)hipFree(victim->second.ptr);
    used -= victim->second.bytes;
    g_tw_cache.erase(victim);
  }
}

int cached_plan(int device, int n, const FftPlan** plan, const double2** tw) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(device, n);
  auto it = g_tw_cache.find(key);
  if (it == g_tw_cache.end()) {
    FftPlan F;
    const int split = fft_split(n, max_fft_len(device));
    if (split < 0) return fail(BLDFM_ELIMIT, "FFT length has no split into two single-workgroup lengths");
    if (split > 0) {  // four-step: the factors are planned on their own
      memset(&F, 0, sizeof(F));
      F.n = n;
      F.n1 = split;
      F.n2 = n / split;
    } else if (!factorize(n, &F)) {
      return fail(BLDFM_ELIMIT, "too many FFT stages");
    }
    F.tw_off = 0;
    std::vector<char> h(kPlanHdr + sizeof(double2) * n, 0);
    memcpy(h.data(), &F, sizeof(F));
    twiddles(n, (double2*)(h.data() + kPlanHdr));
    tw_evict(device, h.size(), key);
    char* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, h.size());
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(fft plan)");
    e = hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(d); return hip_fail(e, "hipMemcpy(fft plan)"); }
    it = g_tw_cache.emplace(key, TwBlock{d, h.size(), 0}).first;
  }
  it->second.last_use = ++g_tw_clock;
  *plan = (const FftPlan*)it->second.ptr;
  *tw = (const double2*)(it->second.ptr + kPlanHdr);
  return 0;
}

void set_lds_attr(int bytes) { bldfm_set_fft_lds_attr(bytes); }

}  // namespace

// ------------------------------------------------------------------- plan
struct bldfm_plan {
  int device = 0;
  int ninst = 0;
  std::vector<InstDesc> desc;
  void* dmeta = nullptr;  // device metadata blob
  PlanDev dev{};
  size_t ws_bytes = 0;
  int64_t tab_doubles = 0, ws_cplx = 0, prof_doubles = 0, src_doubles = 0, out_elems = 0;
  // phase timing (bldfm_plan_timing_begin/end): 7 events per execution
  std::vector<hipEvent_t> events;
  int timing_max = 0, timing_count = 0;
  // surrogate: instance arguments and met setups for the hull check, gate
  // state offsets (double2 units into the complex workspace)
  std::vector<bldfm_instance> inst;
  std::vector<ProfSetup> setups;
  int sur_any = 0;
  int64_t off_gate = 0, off_mode = 0;
  void* sur_ok_dev = nullptr;
  bldfm_sur_hull hull{};  // hull of the surrogate set last (for bldfm_plan_update_met)
};

// Network inputs of one met setup (layout of SurMet); n is the level index
// of the measurement height.
static SurMet sur_met_of(const ProfSetup& s, int n) {
  SurMet m;
  memset(&m, 0, sizeof(m));
  const double il = 1.0 / s.mol;
  double dir = atan2(-s.um, -s.vm);  // compute_wind_fields: u = -U sin(dir), v = -U cos(dir)
  if (dir < 0.0) dir += 2.0 * BLDFM_PI;
  const double wx = s.um / s.absum, wy = s.vm / s.absum;
  m.met[0] = (float)s.ustar;
  m.met[1] = (float)il;
  m.met[2] = (float)log(s.z0);
  m.met[3] = (float)s.zm;
  m.met[4] = (float)s.absum;
  m.met[5] = (float)dir;
  m.met[6] = (float)(s.zm * il);
  m.met[7] = (float)log(s.z0 / s.zm);
  m.wx = (float)wx;
  m.wy = (float)wy;
  m.zm = (float)s.zm;
  // WKB table: interval heights and along-wind speed over diffusivity at the midpoints
  const int nw = std::max(0, std::min(std::min(n, s.nlev - 1), SUR_NWKB));
  m.nw = (float)nw;
  double a[6], b[6];
  for (int j = 0; j < nw; ++j) {
    prof_point<false>(s, j, a);  // float features: libm suffices
    prof_point<false>(s, j + 1, b);
    const double ua = a[1] * wx + a[2] * wy, ub = b[1] * wx + b[2] * wy;
    m.wkb[j][0] = (float)(b[0] - a[0]);
    m.wkb[j][1] = (float)((ua + ub) / (a[5] + b[5]));
  }
  return m;
}

static void destroy_events(bldfm_plan* P) {
  for (hipEvent_t e : P->events) (void)hipEventDestroy(e);
  P->events.clear();
  P->timing_max = P->timing_count = 0;
}

extern "C" {

int bldfm_abi_version(void) { return BLDFM_ABI_VERSION; }

const char* bldfm_last_error(void) { return g_err.c_str(); }

int bldfm_mem_info(int device, size_t* cache_bytes, size_t* free_bytes, size_t* total_bytes) {
  DeviceGuard guard(device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  if (cache_bytes) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    *cache_bytes = tw_cache_bytes(device);
  }
  if (free_bytes || total_bytes) {
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  }
  return 0;
}

static int met_setup(const bldfm_met* m, ProfSetup* s) {
  if (m->n <= 0) return fail(BLDFM_EINVAL, "n must be positive");
  const int rc =
      m->has_prologue
          ? prof_setup_given(m->n, m->meas_height, m->um, m->vm, m->ustar, m->z0, m->mol, m->prsc, m->closure,
                             m->has_stretch, m->stretch, m->tke, m->grid_aa, m->grid_bb, m->grid_zetamax, s)
          : prof_setup(m->n, m->meas_height, m->um, m->vm, m->has_ustar, m->ustar, m->has_z0, m->z0, m->mol,
                       m->prsc, m->closure, m->has_domain_height, m->domain_height, m->has_stretch, m->stretch,
                       m->tke, s);
  if (rc == -1) return fail(BLDFM_EINVAL, "Either z0 or ustar must be provided.");
  if (rc == -2) return fail(BLDFM_EINVAL, "Invalid closure type");
  if (rc == -3) return fail(BLDFM_EINVAL, "vertical grid is empty");
  if (rc == -4) return fail(BLDFM_EINVAL, "arange: cannot compute length");
  if (!std::isfinite(s->z0) || !std::isfinite(s->ustar) || !std::isfinite(s->aa) ||
      !std::isfinite(s->bb) || s->nlev < 2)
    return fail(BLDFM_EINVAL, "profile parameters give a non-finite or degenerate grid");
  return 0;
}

int bldfm_profile_query(const bldfm_met* met, int32_t* nlev, double* z0, double* ustar) {
  if (!met) return fail(BLDFM_EINVAL, "null met");
  ProfSetup s;
  const int rc = met_setup(met, &s);
  if (rc) return rc;
  if (nlev) *nlev = s.nlev;
  if (z0) *z0 = s.z0;
  if (ustar) *ustar = s.ustar;
  return 0;
}

int bldfm_profiles(const bldfm_met* met, double* prof6_dev, hipStream_t stream) {
  if (!met || !prof6_dev) return fail(BLDFM_EINVAL, "null argument");
  ProfSetup s;
  const int rc = met_setup(met, &s);
  if (rc) return rc;
  const int lr = bldfm_launch_profiles(s, prof6_dev, stream);
  if (lr) return fail(BLDFM_EHIP, std::string("profile kernel launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(lr + 1000))));
  return 0;
}

int bldfm_plan_create(int device, const bldfm_instance* inst, int32_t ninst, const int32_t* levels,
                      bldfm_plan** out) {
  return bldfm_plan_create_tuned(device, inst, ninst, levels, nullptr, out);
}

int bldfm_plan_create_tuned(int device, const bldfm_instance* inst, int32_t ninst, const int32_t* levels,
                            const bldfm_tuning* tuning, bldfm_plan** out) {
  if (!inst || ninst <= 0 || !levels || !out) return fail(BLDFM_EINVAL, "bad plan arguments");
  *out = nullptr;
  const int shoot_block = tuning && tuning->shoot_block ? tuning->shoot_block : 256;
  const int xlayout = tuning ? tuning->xpass_layout : 0;
  if (shoot_block != 64 && shoot_block != 128 && shoot_block != 256)
    return fail(BLDFM_EINVAL, "shoot_block must be 64, 128 or 256");
  if (xlayout < 0 || xlayout > 2) return fail(BLDFM_EINVAL, "xpass_layout must be 0, 1 or 2");
  DeviceGuard guard(device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");

  auto* P = new bldfm_plan();
  P->device = device;
  P->ninst = ninst;
  P->desc.resize(ninst);
  std::vector<ProfSetup> met(ninst);
  std::vector<int32_t> met_flag(ninst, 0);
  std::vector<int32_t> ints;
  std::vector<int> plan_n;  // distinct transform lengths
  auto plan_id = [&](int n) {
    for (size_t i = 0; i < plan_n.size(); ++i)
      if (plan_n[i] == n) return (int)i;
    plan_n.push_back(n);
    return (int)plan_n.size() - 1;
  };
  // tab, fwd_rows, fwd_cols, shoot, nyq, inv_cols, inv_rows, sur, spot, transpose,
  // then the four-step steps A, B of fwd_rows, fwd_cols, inv_cols, inv_rows
  constexpr int kKinds = 18;
  std::vector<int64_t> tiles[kKinds];
  for (auto& v : tiles) v.assign(ninst + 1, 0);
  std::vector<SurMet> surmet(ninst);
  P->inst.assign(inst, inst + ninst);
  int64_t off_prof = 0, off_tab = 0, off_src = 0, off_ws = 0, off_out = 0;
  const int32_t* lv = levels;
  int out_f32 = -1, any_disp = 0, variant_used[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  int max_n = 1;

  auto bad = [&](int code, const std::string& m) {
    if (P->dmeta) (void)hipFree(P->dmeta);
    delete P;
    return fail(code, m);
  };

  for (int i = 0; i < ninst; ++i) {
    const bldfm_instance& I = inst[i];
    InstDesc& D = P->desc[i];
    memset(&D, 0, sizeof(D));
    if (I.nx <= 0 || I.ny <= 0) return bad(BLDFM_EINVAL, "grid must be non-empty");
    if (I.nlx <= 0 || I.nly <= 0) return bad(BLDFM_EINVAL, "modes must be positive");
    if (I.px < 0 || I.py < 0) return bad(BLDFM_EINVAL, "halo pad widths must be non-negative");
    if (I.nslot <= 0) return bad(BLDFM_EINVAL, "at least one output level is required");
    D.nx = I.nx; D.ny = I.ny; D.nlx = I.nlx; D.nly = I.nly; D.px = I.px; D.py = I.py;
    D.nxe = I.nx + 2 * I.px;
    D.nye = I.ny + 2 * I.py;
    if (D.nlx > D.nxe || D.nly > D.nye) return bad(BLDFM_EINVAL, "modes exceed the padded grid");
    // Transform lengths: the reference's untruncation pads the nl retained
    // modes by (ne - nl)//2 on each side (src/bldfm/solver.py:265-278), so with
    // an odd ne - nl a footprint's inverse transform has ne - 1 points, and its
    // crop [p, ne - p) keeps min(n, ne - 1 - p) of them; a dispersion solve
    // fails there in the reference (its truncated source spectrum carries
    // nl + 1 modes, solver.py:138-145), so it is refused.
    const int Lx = D.nlx + 2 * ((D.nxe - D.nlx) / 2), Ly = D.nly + 2 * ((D.nye - D.nly) / 2);
    if (!I.footprint && (Lx != D.nxe || Ly != D.nye))
      return bad(BLDFM_EINVAL, "padded grid and mode counts must have equal parity.");
    D.nx = std::min(I.nx, Lx - I.px);  // output columns and rows (n unless ne - nl is odd)
    D.ny = std::min(I.ny, Ly - I.py);
    if (D.nx <= 0 || D.ny <= 0) return bad(BLDFM_EINVAL, "the crop of the inverse transform is empty");
    int split_x = 0, split_y = 0;
    int rc = check_fft_len(Lx, device, "padded x length", &split_x);
    if (!rc) rc = check_fft_len(Ly, device, "padded y length", &split_y);
    if (rc) { delete P; return rc; }
    D.long_x = split_x > 0;
    D.long_y = split_y > 0;
    max_n = std::max(max_n, std::max(lds_len(Lx, split_x), lds_len(Ly, split_y)));
    D.nxb = band_len(D.nlx, Lx);
    D.nyb = band_len(D.nly, Ly);
    if (I.prof_from_met) {
      rc = met_setup(&I.met, &met[i]);
      if (rc) { delete P; return rc; }
      met_flag[i] = 1;
      D.nz = met[i].nlev;
    } else {
      D.nz = I.nz;
    }
    if (D.nz < 2) return bad(BLDFM_EINVAL, "at least two vertical points are required");
    D.nslot = I.nslot;
    D.footprint = I.footprint ? 1 : 0;
    D.analytic = I.analytic ? 1 : 0;
    D.single = I.single ? 1 : 0;
    D.has_shift = I.has_shift ? 1 : 0;
    D.out_f32 = I.out_f32 ? 1 : 0;
    if (out_f32 < 0) out_f32 = D.out_f32;
    if (out_f32 != D.out_f32) return bad(BLDFM_EINVAL, "all instances of a plan must share out_f32");
    if (!D.footprint) any_disp = 1;
    D.plan_x = plan_id(Lx);
    D.plan_y = plan_id(Ly);
    if (split_x) { plan_id(split_x); plan_id(Lx / split_x); }
    if (split_y) { plan_id(split_y); plan_id(Ly / split_y); }
    D.lanes = shoot_lanes(D.nlx, D.nly);
    D.dx = I.dx; D.dy = I.dy;
    if (!(I.dx > 0.0) || !(I.dy > 0.0)) return bad(BLDFM_EINVAL, "grid increments must be positive");
    D.sx = 2.0 * BLDFM_PI / I.dx / D.nxe;
    D.sy = 2.0 * BLDFM_PI / I.dy / D.nye;
    D.vx = 1.0 / (D.nlx * (1.0 / D.nlx));
    D.vy = 1.0 / (D.nly * (1.0 / D.nly));
    D.shx = I.shift_x; D.shy = I.shift_y;
    D.p000 = I.srf_bg_conc;
    // q^0 of the footprint's delta source: the reference divides a complex
    // array by the two lengths (src/bldfm/solver.py:134), and numpy's complex
    // division by a real-valued divisor is Smith's method, a product with the
    // rounded reciprocal: (1/nxe) (1/nye), which differs from (1/nxe)/nye by
    // one ulp for some lengths (960 x 384) -- and the shooting's cancellation
    // turns one ulp of q^0 into 1e-8 of the weakest modes.
    D.q00 = (1.0 / D.nxe) * (1.0 / D.nye);

    // levels: slot map then level index per slot
    D.off_slot = (int64_t)ints.size();
    int nfilled = 0, max_level = -1;
    for (int k = 0; k < 4; ++k) D.snap_level[k] = -1;
    for (int z = 0; z < D.nz; ++z) {
      const int s = lv[z];
      if (s >= 0) {
        if (s != nfilled || s >= D.nslot) return bad(BLDFM_EINVAL, "level->slot map must number slots 0,1,.. in ascending level order");
        if (s < 4) D.snap_level[s] = z;
        ++nfilled;
        max_level = z;
      }
      ints.push_back(s);
    }
    for (int s = 0; s < D.nslot; ++s) {
      const int li = lv[D.nz + s];
      if (D.analytic && (li < 0 || li >= D.nz)) return bad(BLDFM_EINVAL, "analytic level index out of range");
      ints.push_back(li);
    }
    lv += D.nz + D.nslot;
    D.nfilled = nfilled;
    D.max_level = max_level;
    D.variant = D.analytic ? 3 : (nfilled <= 1 ? 0 : (nfilled <= 4 ? 1 : 2));
    D.exact = I.fast ? 0 : 1;
    variant_used[D.variant][D.exact] = 1;
    const bool want_sur = I.surrogate != 0 && I.prof_from_met && D.footprint;
    if (want_sur) {
      surmet[i] = sur_met_of(met[i], I.met.n);
      P->sur_any = 1;
    } else {
      memset(&surmet[i], 0, sizeof(SurMet));
    }

    // buffer offsets
    D.off_prof = off_prof; off_prof += 6LL * D.nz;
    D.off_tab = off_tab; off_tab += (int64_t)D.nz * BLDFM_LEVEL_STRIDE;
    D.off_src = off_src;
    if (!D.footprint) off_src += (int64_t)D.ny * D.nx;
    if (!D.footprint) {
      D.off_aw = off_ws; off_ws += (int64_t)D.ny * D.nlx;
      D.off_q0 = off_ws; off_ws += (int64_t)D.nlx * D.nly;
    } else {
      D.off_aw = D.off_q0 = -1;
    }
    D.off_z = off_ws; off_ws += (int64_t)D.nslot * D.nxb * D.nyb;
    D.off_nyq = off_ws; off_ws += (int64_t)D.nslot * (D.nlx + D.nly) * 2;
    D.off_v = off_ws; off_ws += (int64_t)D.nslot * D.nxb * D.ny;
    if (xlayout == 2) { D.off_vt = off_ws; off_ws += (int64_t)D.nslot * D.nxb * D.ny; }
    D.off_out = off_out; off_out += (int64_t)D.nslot * D.ny * D.nx;
    // four-step scratch: lines x n of the largest split pass
    int64_t long_elems = 0;
    if (D.long_x) {
      long_elems = std::max(long_elems, (int64_t)D.nslot * D.ny * Lx);
      if (!D.footprint) long_elems = std::max(long_elems, (int64_t)D.ny * Lx);
    }
    if (D.long_y) {
      long_elems = std::max(long_elems, (int64_t)D.nslot * D.nxb * Ly);
      if (!D.footprint) long_elems = std::max(long_elems, (int64_t)D.nlx * Ly);
    }
    D.off_long = long_elems ? off_ws : -1;
    off_ws += long_elems;

    // tiles: tab, fwd_rows, fwd_cols, shoot, nyq, inv_cols, inv_rows
    const int fixc = nyq_fix_count(D.nlx, D.nly, D.nxb, D.nyb);
    const int64_t nspot = (D.lanes + BLDFM_SUR_SPOT_STRIDE - 1) / BLDFM_SUR_SPOT_STRIDE;
    // lines of each pass; a split axis moves them to the four-step tiles
    const int64_t lines_fr = D.footprint ? 0 : D.ny, lines_fc = D.footprint ? 0 : D.nlx;
    const int64_t lines_ic = (int64_t)D.nslot * D.nxb, lines_ir = (int64_t)D.nslot * D.ny;
    const int64_t x1 = split_x, x2 = split_x ? Lx / split_x : 0;
    const int64_t y1 = split_y, y2 = split_y ? Ly / split_y : 0;
    const int64_t nt[kKinds] = {
        (D.nz + 255) / 256,
        D.long_x ? 0 : lines_fr,
        D.long_y ? 0 : lines_fc,
        (D.lanes + shoot_block - 1) / shoot_block,
        fixc > 0 ? (int64_t)D.nslot * ((fixc + 255) / 256) : 0,
        D.long_y ? 0 : lines_ic,
        D.long_x ? 0 : lines_ir,
        want_sur ? (D.lanes + SUR_PTS - 1) / SUR_PTS : 0,
        want_sur ? (nspot + 255) / 256 : 0,
        xlayout == 2 ? (int64_t)D.nslot * ((D.nxb + 31) / 32) * ((D.ny + 31) / 32) : 0,
        // four-step steps A (lines x n2) and B (lines x n1) per pass
        lines_fr * x2, lines_fr * x1, lines_fc * y2, lines_fc * y1,
        lines_ic * y2, lines_ic * y1, lines_ir * x2, lines_ir * x1,
    };
    for (int k = 10; k < kKinds; ++k)
      if (nt[k] >= ((int64_t)1 << 31)) return bad(BLDFM_ELIMIT, "too many four-step transforms in one instance");
    for (int k = 0; k < kKinds; ++k) tiles[k][i + 1] = tiles[k][i] + nt[k];
  }
  if (lv - levels > (int64_t)1 << 31) return bad(BLDFM_EINVAL, "level table too large");
  P->setups = met;
  if (P->sur_any) {  // gate sums (double2) and modes (int32) per instance
    P->off_gate = off_ws;
    off_ws += ninst;
    P->off_mode = off_ws;
    off_ws += (ninst + 3) / 4;
  }

  // launch shape of the FFT passes (validated before any device allocation)
  const int fft_small = max_n <= BLDFM_FFT_SMALL_N ? 1 : 0;
  int fft_threads = fft_threads_for(max_n);
  if (tuning && tuning->fft_threads) {
    const int cap = fft_small ? 256 : 1024;  // launch bounds of the two FFT kernel families
    const int t = tuning->fft_threads;
    if (t % 64 != 0 || t < fft_threads || t > cap)
      return bad(BLDFM_EINVAL, "fft_threads must be a multiple of 64 between " + std::to_string(fft_threads) +
                                   " and " + std::to_string(cap) + " for this plan");
    fft_threads = t;
  }

  // FFT plans and twiddles
  std::vector<FftPlan> plans(plan_n.size());
  int64_t tw_total = 0;
  const int fft_limit = max_fft_len(device);
  auto id_of = [&](int n) {
    for (size_t i = 0; i < plan_n.size(); ++i)
      if (plan_n[i] == n) return (int)i;
    return -1;
  };
  for (size_t k = 0; k < plan_n.size(); ++k) {
    const int n = plan_n[k];
    const int split = fft_split(n, fft_limit);
    if (split > 0) {
      memset(&plans[k], 0, sizeof(FftPlan));
      plans[k].n = n;
      plans[k].n1 = split;
      plans[k].n2 = n / split;
      plans[k].plan1 = id_of(split);
      plans[k].plan2 = id_of(n / split);
    } else if (!factorize(n, &plans[k])) {
      return bad(BLDFM_ELIMIT, "too many FFT stages");
    }
    plans[k].tw_off = tw_total;
    tw_total += n;
  }
  std::vector<double2> tw(tw_total);
  for (size_t k = 0; k < plan_n.size(); ++k) twiddles(plan_n[k], tw.data() + plans[k].tw_off);

  // device metadata blob
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off = (size_t)align_up((int64_t)(off + bytes), 256); return o; };
  const size_t o_desc = take(sizeof(InstDesc) * ninst);
  const size_t o_met = take(sizeof(ProfSetup) * ninst);
  const size_t o_flag = take(sizeof(int32_t) * ninst);
  const size_t o_ints = take(sizeof(int32_t) * std::max<size_t>(ints.size(), 1));
  const size_t o_plans = take(sizeof(FftPlan) * plans.size());
  const size_t o_tw = take(sizeof(double2) * tw_total);
  const size_t o_surmet = take(sizeof(SurMet) * ninst);
  // tile records {instance, local index} of each kernel
  std::vector<TileRec> recs[kKinds];
  for (int k = 0; k < kKinds; ++k) {
    if (tiles[k][ninst] > ((int64_t)1 << 31)) return bad(BLDFM_ELIMIT, "too many tiles in one plan");
    recs[k].resize((size_t)tiles[k][ninst]);
    for (int i = 0; i < ninst; ++i)
      for (int64_t t = tiles[k][i]; t < tiles[k][i + 1]; ++t) recs[k][t] = TileRec{i, (int32_t)(t - tiles[k][i])};
  }
  size_t o_tiles[kKinds];
  for (int k = 0; k < kKinds; ++k) o_tiles[k] = take(sizeof(TileRec) * std::max<size_t>(recs[k].size(), 1));
  std::vector<char> blob(off, 0);
  memcpy(blob.data() + o_desc, P->desc.data(), sizeof(InstDesc) * ninst);
  memcpy(blob.data() + o_met, met.data(), sizeof(ProfSetup) * ninst);
  memcpy(blob.data() + o_flag, met_flag.data(), sizeof(int32_t) * ninst);
  if (!ints.empty()) memcpy(blob.data() + o_ints, ints.data(), sizeof(int32_t) * ints.size());
  memcpy(blob.data() + o_plans, plans.data(), sizeof(FftPlan) * plans.size());
  memcpy(blob.data() + o_tw, tw.data(), sizeof(double2) * tw_total);
  memcpy(blob.data() + o_surmet, surmet.data(), sizeof(SurMet) * ninst);
  for (int k = 0; k < kKinds; ++k)
    if (!recs[k].empty()) memcpy(blob.data() + o_tiles[k], recs[k].data(), sizeof(TileRec) * recs[k].size());
  hipError_t e = hipMalloc(&P->dmeta, off);
  if (e != hipSuccess) { P->dmeta = nullptr; delete P; return hip_fail(e, "hipMalloc(plan metadata)"); }
  e = hipMemcpy(P->dmeta, blob.data(), off, hipMemcpyHostToDevice);
  if (e != hipSuccess) { (void)hipFree(P->dmeta); delete P; return hip_fail(e, "hipMemcpy(plan metadata)"); }

  char* base = (char*)P->dmeta;
  PlanDev& V = P->dev;
  V.desc = (const InstDesc*)(base + o_desc);
  V.met = (const ProfSetup*)(base + o_met);
  V.met_flag = (const int32_t*)(base + o_flag);
  V.ints = (const int32_t*)(base + o_ints);
  V.plans = (const FftPlan*)(base + o_plans);
  V.tw = (const double2*)(base + o_tw);
  V.t_tab = (const TileRec*)(base + o_tiles[0]);
  V.t_fwd_rows = (const TileRec*)(base + o_tiles[1]);
  V.t_fwd_cols = (const TileRec*)(base + o_tiles[2]);
  V.t_shoot = (const TileRec*)(base + o_tiles[3]);
  V.t_nyq = (const TileRec*)(base + o_tiles[4]);
  V.t_inv_cols = (const TileRec*)(base + o_tiles[5]);
  V.t_inv_rows = (const TileRec*)(base + o_tiles[6]);
  V.t_sur = (const TileRec*)(base + o_tiles[7]);
  V.t_spot = (const TileRec*)(base + o_tiles[8]);
  V.t_tr = (const TileRec*)(base + o_tiles[9]);
  for (int k = 0; k < 8; ++k) {
    V.t_long[k] = (const TileRec*)(base + o_tiles[10 + k]);
    V.tot_long[k] = tiles[10 + k][ninst];
  }
  V.tot_tr = tiles[9][ninst];
  V.shoot_block = shoot_block;
  V.xlayout = xlayout;
  V.sur_met = (const SurMet*)(base + o_surmet);
  V.sur_ok = nullptr;
  V.sur_net = nullptr;
  V.sur_tol = 0.0;
  V.sur_stride = BLDFM_SUR_SPOT_STRIDE;
  V.tot_sur = tiles[7][ninst];
  V.tot_spot = tiles[8][ninst];
  V.ninst = ninst;
  V.any_dispersion = any_disp;
  for (int k = 0; k < 4; ++k)
    for (int e = 0; e < 2; ++e) V.variant_used[k][e] = variant_used[k][e];
  V.fft_small = fft_small;
  V.fft_threads = fft_threads;
  V.fft_lds_bytes = fft_lds_bytes(max_n);  // non-decreasing in n
  V.tot_tab = tiles[0][ninst];
  V.tot_fwd_rows = tiles[1][ninst];
  V.tot_fwd_cols = tiles[2][ninst];
  V.tot_shoot = tiles[3][ninst];
  V.tot_nyq = tiles[4][ninst];
  V.tot_inv_cols = tiles[5][ninst];
  V.tot_inv_rows = tiles[6][ninst];
  set_lds_attr(V.fft_lds_bytes);

  P->tab_doubles = off_tab;
  P->ws_cplx = off_ws;
  P->prof_doubles = off_prof;
  P->src_doubles = off_src;
  P->out_elems = off_out;
  P->ws_bytes = (size_t)align_up(off_tab * 8, 256) + (size_t)off_ws * 16;
  *out = P;
  return 0;
}

int bldfm_plan_destroy(bldfm_plan* plan) {
  if (!plan) return 0;
  if (plan->dmeta || plan->sur_ok_dev || !plan->events.empty()) {
    DeviceGuard g(plan->device);
    destroy_events(plan);
    if (plan->dmeta) (void)hipFree(plan->dmeta);
    if (plan->sur_ok_dev) (void)hipFree(plan->sur_ok_dev);
  }
  delete plan;
  return 0;
}

int bldfm_plan_sizes(const bldfm_plan* P, size_t* ws, int64_t* prof, int64_t* src, int64_t* out) {
  if (!P) return fail(BLDFM_EINVAL, "null plan");
  if (ws) *ws = P->ws_bytes;
  if (prof) *prof = P->prof_doubles;
  if (src) *src = P->src_doubles;
  if (out) *out = P->out_elems;
  return 0;
}

int bldfm_plan_offsets(const bldfm_plan* P, int32_t i, int64_t* off_prof, int64_t* off_src,
                       int64_t* off_out, int32_t* nz) {
  if (!P || i < 0 || i >= P->ninst) return fail(BLDFM_EINVAL, "bad plan or instance index");
  const InstDesc& D = P->desc[i];
  if (off_prof) *off_prof = D.off_prof;
  if (off_src) *off_src = D.off_src;
  if (off_out) *off_out = D.off_out;
  if (nz) *nz = D.nz;
  return 0;
}

}  // extern "C"

// Workspace of bldfm_plan_execute_static: one array per GPU the library runs
// on, held until the process ends.  Host code never uses the array's name as
// an address: the device address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device.
static __device__ __align__(256) char scratch[BLDFM_STATIC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> address of that device's `scratch`

// Address of the current device's scratch array (device `dev` must be current).
int scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    it = g_scratch_ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return 0;
}
}  // namespace

extern "C" {

int bldfm_plan_execute_static(bldfm_plan* P, double* prof, const double* src, void* conc, void* flx,
                              hipStream_t stream) {
  if (!P) return fail(BLDFM_EINVAL, "null plan");
  if (!prof || !conc || !flx) return fail(BLDFM_EINVAL, "null buffer");
  if (P->ws_bytes > (size_t)BLDFM_STATIC_SCRATCH_BYTES)
    return fail(BLDFM_ELIMIT, "plan workspace of " + std::to_string(P->ws_bytes) +
                                  " bytes exceeds the static scratch of " +
                                  std::to_string((size_t)BLDFM_STATIC_SCRATCH_BYTES) +
                                  " bytes; use bldfm_plan_execute with a workspace");
  void* ws = nullptr;
  {
    DeviceGuard guard(P->device);
    if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
    const int rc = scratch_address(P->device, &ws);
    if (rc) return rc;
  }
  return bldfm_plan_execute(P, prof, src, conc, flx, ws, stream);
}

int bldfm_plan_execute(bldfm_plan* P, double* prof, const double* src, void* conc, void* flx,
                       void* workspace, hipStream_t stream) {
  if (!P) return fail(BLDFM_EINVAL, "null plan");
  if (!prof || !conc || !flx || !workspace) return fail(BLDFM_EINVAL, "null buffer");
  if (P->dev.any_dispersion && !src) return fail(BLDFM_EINVAL, "dispersion instances need srf_flx");
  DeviceGuard guard(P->device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  ExecBufs B;
  B.prof = prof;
  B.src = src;
  B.conc = conc;
  B.flx = flx;
  B.tab = (double*)workspace;
  B.ws = (double2*)((char*)workspace + align_up(P->tab_doubles * 8, 256));
  const bool sur = P->sur_any && P->dev.sur_net != nullptr;
  B.gate = sur ? B.ws + P->off_gate : nullptr;
  B.mode = sur ? (int32_t*)(B.ws + P->off_mode) : nullptr;
  hipEvent_t* ev = nullptr;
  if (P->timing_count < P->timing_max) ev = P->events.data() + (BLDFM_PHASES + 1) * (P->timing_count++);
  const int rc = bldfm_launch_all(P->dev, B, kGridCap, stream, ev);
  if (rc == -1999) return fail(BLDFM_EHIP, "hipEventRecord failed");
  if (rc) return fail(BLDFM_EHIP, std::string("kernel launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(rc + 1000))));
  return 0;
}

int bldfm_plan_timing_begin(bldfm_plan* P, int32_t max_execs) {
  if (!P || max_execs <= 0 || max_execs > 100000) return fail(BLDFM_EINVAL, "bad timing arguments");
  DeviceGuard guard(P->device);
  destroy_events(P);
  P->events.resize((size_t)(BLDFM_PHASES + 1) * max_execs, nullptr);
  for (auto& e : P->events) {
    hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) { destroy_events(P); return hip_fail(err, "hipEventCreate"); }
  }
  P->timing_max = max_execs;
  P->timing_count = 0;
  return 0;
}

int bldfm_plan_timing_end(bldfm_plan* P, double* phase_ms, int32_t* nexecs) {
  if (!P || !phase_ms) return fail(BLDFM_EINVAL, "bad timing arguments");
  DeviceGuard guard(P->device);
  for (int k = 0; k < BLDFM_PHASES; ++k) phase_ms[k] = 0.0;
  const int n = P->timing_count;
  int rc = 0;
  if (n > 0) {
    hipError_t err = hipEventSynchronize(P->events[(BLDFM_PHASES + 1) * (n - 1) + BLDFM_PHASES]);
    if (err != hipSuccess) rc = hip_fail(err, "hipEventSynchronize");
    for (int x = 0; x < n && !rc; ++x)
      for (int k = 0; k < BLDFM_PHASES; ++k) {
        float ms = 0.f;
        const int e0 = (BLDFM_PHASES + 1) * x + k;
        err = hipEventElapsedTime(&ms, P->events[e0], P->events[e0 + 1]);
        if (err != hipSuccess) { rc = hip_fail(err, "hipEventElapsedTime"); break; }
        phase_ms[k] += ms;
      }
  }
  if (nexecs) *nexecs = n;
  destroy_events(P);
  return rc;
}

int bldfm_stability(const double* x, int64_t n, double* psi, double* phi, hipStream_t stream) {
  if (n < 0 || (n > 0 && !x)) return fail(BLDFM_EINVAL, "bad stability arguments");
  const int rc = bldfm_launch_stability(x, n, psi, phi, stream);
  if (rc) return fail(BLDFM_EHIP, std::string("stability kernel launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(rc + 1000))));
  return 0;
}

int bldfm_wind_fields(const double* speed, const double* dir, int64_t n, double* u, double* v,
                      hipStream_t stream) {
  if (n < 0 || (n > 0 && (!speed || !dir || !u || !v))) return fail(BLDFM_EINVAL, "bad wind arguments");
  const int rc = bldfm_launch_wind(speed, dir, n, u, v, stream);
  if (rc) return fail(BLDFM_EHIP, std::string("wind kernel launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(rc + 1000))));
  return 0;
}

int bldfm_math_probe(int32_t op, const double* a, const double* b, int64_t n, double* out,
                     hipStream_t stream) {
  if (op < 0 || op > 5 || n < 0 || (n > 0 && (!a || !out || (op != 5 && !b))))
    return fail(BLDFM_EINVAL, "bad math probe arguments");
  const int rc = bldfm_launch_math_probe(op, a, b, n, out, stream);
  if (rc) return fail(BLDFM_EHIP, std::string("math probe launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(rc + 1000))));
  return 0;
}

size_t bldfm_fft2_workspace_bytes(int32_t n0, int32_t n1, int32_t batch) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int limit = max_fft_len(dev);
  const bool split = fft_split(n0, limit) != 0 || fft_split(n1, limit) != 0;
  return (size_t)n0 * n1 * batch * sizeof(double2) * (split ? 2 : 1);
}

int bldfm_fft2(int device, int32_t n0, int32_t n1, int32_t batch, const double* in, double* out,
               int32_t sign, double scale, void* workspace, hipStream_t stream) {
  if (n0 <= 0 || n1 <= 0 || batch <= 0 || !in || !out || !workspace)
    return fail(BLDFM_EINVAL, "bad fft2 arguments");
  if (sign != 1 && sign != -1) return fail(BLDFM_EINVAL, "sign must be +1 or -1");
  DeviceGuard guard(device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  int s0 = 0, s1 = 0;
  int rc = check_fft_len(n0, device, "fft2 axis 0 length", &s0);
  if (!rc) rc = check_fft_len(n1, device, "fft2 axis 1 length", &s1);
  if (rc) return rc;
  FftSplit A[2];
  const int lens[2] = {n0, n1}, splits[2] = {s0, s1};
  for (int a = 0; a < 2; ++a) {
    memset(&A[a], 0, sizeof(FftSplit));
    rc = cached_plan(device, lens[a], &A[a].F, &A[a].tw);
    if (!rc && splits[a]) {
      A[a].n1 = splits[a];
      A[a].n2 = lens[a] / splits[a];
      rc = cached_plan(device, A[a].n1, &A[a].F1, &A[a].tw1);
      if (!rc) rc = cached_plan(device, A[a].n2, &A[a].F2, &A[a].tw2);
    }
    if (rc) return rc;
  }
  {  // a later cached_plan may have evicted an earlier block of this call: look them up again
    std::lock_guard<std::mutex> lk(g_tw_mu);
    for (int a = 0; a < 2; ++a) {
      const int ns[3] = {lens[a], A[a].n1, A[a].n2};
      for (int q = 0; q < (splits[a] ? 3 : 1); ++q)
        if (g_tw_cache.find(std::make_pair(device, ns[q])) == g_tw_cache.end())
          return fail(BLDFM_ELIMIT, "fft2 plans exceed the twiddle cache");
    }
  }
  const int nmax = std::max(lds_len(n0, s0), lds_len(n1, s1));
  set_lds_attr(fft_lds_bytes(nmax));
  double2* mid = (double2*)workspace;
  double2* scr = (s0 || s1) ? mid + (int64_t)n0 * n1 * batch : nullptr;
  rc = bldfm_launch_fft2(A[0], n0, A[1], n1, batch, (const double2*)in, mid, scr, (double2*)out, sign, scale,
                         fft_threads_for(nmax), fft_lds_bytes(nmax), kGridCap, stream);
  if (rc) return fail(BLDFM_EHIP, std::string("fft2 launch failed: ") +
                                      hipGetErrorString((hipError_t)(-(rc + 1000))));
  return 0;
}

// ------------------------------------------------------------------ surrogate
static int launch_fail(int rc, const char* what) {
  return fail(BLDFM_EHIP, std::string(what) + " launch failed: " +
                              hipGetErrorString((hipError_t)(-(rc + 1000))));
}

static constexpr int64_t kSurTrainBytesPerPoint =
    (int64_t)(SUR_IN + SUR_NL * SUR_W) * 2 + 2LL * SUR_NL * SUR_W * 2 + 8;

int bldfm_sur_sizes(int64_t* nparam, int64_t* net_bytes, int64_t* train_bytes_per_point,
                    int32_t* block_points) {
  if (nparam) *nparam = SUR_NPARAM;
  if (net_bytes) *net_bytes = SUR_NET_BYTES;
  if (train_bytes_per_point) *train_bytes_per_point = kSurTrainBytesPerPoint;
  if (block_points) *block_points = SUR_PTS;
  return 0;
}

int bldfm_sur_met(const bldfm_met* met, int32_t nmet, float* surmet) {
  if (!met || nmet <= 0 || !surmet) return fail(BLDFM_EINVAL, "bad sur_met arguments");
  for (int i = 0; i < nmet; ++i) {
    ProfSetup s;
    const int rc = met_setup(met + i, &s);
    if (rc) return rc;
    const SurMet m = sur_met_of(s, met[i].n);
    memcpy(surmet + (size_t)i * (sizeof(SurMet) / sizeof(float)), &m, sizeof(SurMet));
  }
  return 0;
}

int bldfm_transfer_tables(int device, const bldfm_met* met, int32_t nmet, int32_t rows, double* tab,
                          int32_t* nz, hipStream_t stream) {
  if (!met || nmet <= 0 || rows < 2 || !tab || !nz) return fail(BLDFM_EINVAL, "bad transfer_tables arguments");
  std::vector<ProfSetup> S(nmet);
  for (int i = 0; i < nmet; ++i) {
    const int rc = met_setup(met + i, &S[i]);
    if (rc) return rc;
    if (S[i].nlev > rows)
      return fail(BLDFM_EINVAL, "met sample " + std::to_string(i) + " has more levels than rows");
    nz[i] = S[i].nlev;
  }
  DeviceGuard guard(device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(ProfSetup) * nmet);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(met setups)");
  e = hipMemcpy(d, S.data(), sizeof(ProfSetup) * nmet, hipMemcpyHostToDevice);
  int rc = 0;
  if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(met setups)");
  if (!rc) {
    const int lr = bldfm_launch_tables((const ProfSetup*)d, nmet, rows, tab, stream);
    if (lr) rc = launch_fail(lr, "level table");
  }
  if (!rc) {
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  }
  (void)hipFree(d);
  return rc;
}

int bldfm_transfer_solve(const double* tab, int32_t rows, const int32_t* nz, const int32_t* lvl,
                         int32_t nmet, const int32_t* blk_met, int32_t nblk, const double* kxy,
                         double* g, int32_t exact, hipStream_t stream) {
  if (!tab || !nz || !lvl || !blk_met || !kxy || !g || rows < 2 || nmet <= 0 || nblk <= 0)
    return fail(BLDFM_EINVAL, "bad transfer_solve arguments");
  const int rc = bldfm_launch_transfer(tab, rows, nz, lvl, nmet, blk_met, nblk, (const double2*)kxy,
                                       (double2*)g, exact ? 1 : 0, stream);
  return rc ? launch_fail(rc, "transfer") : 0;
}

int bldfm_sur_pack(const float* params, void* net, hipStream_t stream) {
  if (!params || !net) return fail(BLDFM_EINVAL, "bad sur_pack arguments");
  const int rc = bldfm_launch_sur_adam(const_cast<float*>(params), nullptr, nullptr, nullptr, (uint8_t*)net, 0,
                                       0.f, 0.f, 0.f, 0.f, 0.f, 0.f, stream);
  return rc ? launch_fail(rc, "surrogate pack") : 0;
}

static int sur_points_ok(int64_t npts) {
  if (npts <= 0 || npts % SUR_PTS) return fail(BLDFM_EINVAL, "npts must be a positive multiple of the block size");
  if (npts / SUR_PTS > ((int64_t)1 << 31) - 1) return fail(BLDFM_ELIMIT, "too many points");
  return 0;
}

int bldfm_sur_forward(const void* net, const float* surmet, int32_t nmet, const int32_t* blk_met,
                      const double* kxy, int64_t npts, float* phi, hipStream_t stream) {
  if (!net || !surmet || nmet <= 0 || !blk_met || !kxy || !phi) return fail(BLDFM_EINVAL, "bad sur_forward arguments");
  int rc = sur_points_ok(npts);
  if (rc) return rc;
  rc = bldfm_launch_sur_forward((const uint8_t*)net, (const SurMet*)surmet, nmet, blk_met, (const double2*)kxy,
                                nullptr, npts, nullptr, nullptr, nullptr, (float2*)phi, nullptr, 0, stream);
  return rc ? launch_fail(rc, "surrogate forward") : 0;
}

int bldfm_sur_train_grad(const void* net, const float* surmet, int32_t nmet, const int32_t* blk_met,
                         const double* kxy, const double* g, int64_t npts, void* ws, float* grad,
                         float* loss, float* phi, hipStream_t stream) {
  if (!net || !surmet || nmet <= 0 || !blk_met || !kxy || !g || !ws || !grad || !loss)
    return fail(BLDFM_EINVAL, "bad sur_train_grad arguments");
  int rc = sur_points_ok(npts);
  if (rc) return rc;
  __bf16* xs = (__bf16*)ws;
  __bf16* zs = xs + (int64_t)(SUR_IN + SUR_NL * SUR_W) * npts;
  __bf16* dzs = zs + (int64_t)SUR_NL * SUR_W * npts;
  float2* dphi = (float2*)(dzs + (int64_t)SUR_NL * SUR_W * npts);
  rc = bldfm_launch_sur_forward((const uint8_t*)net, (const SurMet*)surmet, nmet, blk_met, (const double2*)kxy,
                                (const double2*)g, npts, xs, zs, dphi, (float2*)phi, loss, 1, stream);
  if (rc) return launch_fail(rc, "surrogate forward");
  const int chunk = npts % 2048 == 0 ? 2048 : SUR_PTS;
  rc = bldfm_launch_sur_backward((const uint8_t*)net, npts, xs, zs, dphi, dzs, grad, chunk, stream);
  return rc ? launch_fail(rc, "surrogate backward") : 0;
}

int bldfm_sur_adam(float* params, float* grad, float* m, float* v, void* net, float lr, float beta1,
                   float beta2, float eps, int32_t step, hipStream_t stream) {
  if (!params || !grad || !m || !v || !net || step < 1) return fail(BLDFM_EINVAL, "bad sur_adam arguments");
  const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, step)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, step)));
  const int rc = bldfm_launch_sur_adam(params, grad, m, v, (uint8_t*)net, 1, lr, beta1, beta2, eps, c1, c2, stream);
  return rc ? launch_fail(rc, "surrogate adam") : 0;
}

// Whether instance i of a plan lies inside a network's training hull.
static bool sur_in_hull(const bldfm_plan* P, int i, const bldfm_sur_hull& H) {
  const bldfm_instance& I = P->inst[i];
  const InstDesc& D = P->desc[i];
  if (!I.surrogate || !I.prof_from_met || !D.footprint || D.analytic) return false;
  if (D.nslot != 1 || D.nfilled != 1 || D.max_level != I.met.n) return false;
  const bldfm_met& m = I.met;
  if (m.closure != H.closure || m.n != H.n || m.has_stretch || !m.has_domain_height || m.prsc != 1.0) return false;
  if (!(fabs(m.domain_height / m.meas_height - H.top_ratio) <= 1e-9 * H.top_ratio)) return false;
  const ProfSetup& s = P->setups[i];
  const double il = 1.0 / s.mol;
  if (!(s.ustar >= H.ustar_lo && s.ustar <= H.ustar_hi)) return false;
  if (!(il >= H.invl_lo && il <= H.invl_hi)) return false;
  if (!(s.z0 >= H.z0_lo && s.z0 <= H.z0_hi && s.z0 <= H.z0_zm_max * s.zm)) return false;
  if (!(s.zm >= H.zm_lo && s.zm <= H.zm_hi)) return false;
  const double kx = D.sx * (D.nlx / 2) * D.vx, ky = D.sy * (D.nly / 2) * D.vy;
  if (!(kx <= H.kmax && ky <= H.kmax)) return false;
  // enough spot-checked modes for the gate's error estimate to mean something
  return D.lanes / BLDFM_SUR_SPOT_STRIDE >= BLDFM_SUR_MIN_SPOT;
}

int bldfm_plan_set_surrogate(bldfm_plan* P, const void* net, const bldfm_sur_hull* hull, double tol) {
  if (!P) return fail(BLDFM_EINVAL, "null plan");
  if (!net) {
    P->dev.sur_net = nullptr;
    return 0;
  }
  if (!hull || !(tol > 0.0)) return fail(BLDFM_EINVAL, "bad surrogate hull or tolerance");
  if (!P->sur_any) return fail(BLDFM_EINVAL, "the plan has no surrogate-flagged footprint instance");
  DeviceGuard guard(P->device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  std::vector<int32_t> ok(P->ninst);
  for (int i = 0; i < P->ninst; ++i) ok[i] = sur_in_hull(P, i, *hull) ? 1 : 0;
  hipError_t e = hipSuccess;
  if (!P->sur_ok_dev) e = hipMalloc(&P->sur_ok_dev, sizeof(int32_t) * P->ninst);
  if (e != hipSuccess) { P->sur_ok_dev = nullptr; return hip_fail(e, "hipMalloc(surrogate flags)"); }
  e = hipMemcpy(P->sur_ok_dev, ok.data(), sizeof(int32_t) * P->ninst, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(surrogate flags)");
  P->dev.sur_ok = (const int32_t*)P->sur_ok_dev;
  P->dev.sur_net = (const uint8_t*)net;
  P->dev.sur_tol = tol;
  P->hull = *hull;
  return 0;
}

int bldfm_plan_update_met(bldfm_plan* P, const bldfm_met* mets, int32_t n, hipStream_t stream) {
  if (!P || !mets || n != P->ninst) return fail(BLDFM_EINVAL, "update_met needs one met per plan instance");
  DeviceGuard guard(P->device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  // validate everything before changing anything
  std::vector<ProfSetup> setups = P->setups;
  for (int i = 0; i < n; ++i) {
    if (!P->inst[i].prof_from_met) continue;
    ProfSetup s;
    const int rc = met_setup(&mets[i], &s);
    if (rc) return rc;
    if (s.nlev != P->desc[i].nz)
      return fail(BLDFM_EINVAL, "instance " + std::to_string(i) + ": the new meteorology changes the vertical grid (" +
                                    std::to_string(P->desc[i].nz) + " -> " + std::to_string(s.nlev) +
                                    " levels); build a new plan");
    setups[i] = s;
  }
  std::vector<SurMet> surmet(n);
  for (int i = 0; i < n; ++i) {
    memset(&surmet[i], 0, sizeof(SurMet));
    if (!P->inst[i].prof_from_met) continue;
    P->inst[i].met = mets[i];
    if (P->inst[i].surrogate && P->desc[i].footprint) surmet[i] = sur_met_of(setups[i], mets[i].n);
  }
  P->setups = setups;
  // stream-ordered after earlier executions, which read these arrays
  hipError_t e = hipMemcpyAsync((void*)P->dev.met, setups.data(), sizeof(ProfSetup) * n, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess && P->sur_any)
    e = hipMemcpyAsync((void*)P->dev.sur_met, surmet.data(), sizeof(SurMet) * n, hipMemcpyHostToDevice, stream);
  std::vector<int32_t> ok(n, 0);
  if (e == hipSuccess && P->dev.sur_net && P->sur_ok_dev) {
    for (int i = 0; i < n; ++i) ok[i] = sur_in_hull(P, i, P->hull) ? 1 : 0;
    e = hipMemcpyAsync(P->sur_ok_dev, ok.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the host arrays go out of scope
  if (e != hipSuccess) return hip_fail(e, "uploading new meteorology");
  return 0;
}

int bldfm_plan_sur_stats(const bldfm_plan* P, const void* workspace, double* err, int32_t* mode,
                         hipStream_t stream) {
  if (!P || !workspace) return fail(BLDFM_EINVAL, "bad surrogate stats arguments");
  if (!P->sur_any || !P->dev.sur_net) return fail(BLDFM_EINVAL, "the plan serves no instance by its surrogate");
  DeviceGuard guard(P->device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  const double2* ws = (const double2*)((const char*)workspace + align_up(P->tab_doubles * 8, 256));
  std::vector<double2> gate(P->ninst);
  std::vector<int32_t> md(P->ninst);
  hipError_t e = hipMemcpyAsync(gate.data(), ws + P->off_gate, sizeof(double2) * P->ninst, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(md.data(), ws + P->off_mode, sizeof(int32_t) * P->ninst, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "reading surrogate gate state");
  std::vector<int32_t> ok(P->ninst);
  e = hipMemcpy(ok.data(), P->sur_ok_dev, sizeof(int32_t) * P->ninst, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "reading surrogate flags");
  for (int i = 0; i < P->ninst; ++i) {
    if (err) err[i] = (ok[i] && gate[i].y > 0.0) ? sqrt(gate[i].x / gate[i].y) : NAN;
    if (mode) mode[i] = md[i];
  }
  return 0;
}

// ------------------------------------------------------------------ KM01
// Grid length of numpy's arange(start, stop, step): ceil((stop - start) / step).
static int64_t arange_len(double start, double stop, double step) {
  const double n = std::ceil((stop - start) / step);
  return n > 0 ? (int64_t)n : 0;
}

int bldfm_km01_grid(const bldfm_km01* p, int32_t* nx, int32_t* ny) {
  if (!p || !(p->res > 0.0)) return fail(BLDFM_EINVAL, "grid_res must be positive");
  const int64_t x = arange_len(p->xmin + 0.5 * p->res, p->xmax, p->res);
  const int64_t y = arange_len(p->ymax - 0.5 * p->res, p->ymin, -p->res);
  if (x > INT32_MAX || y > INT32_MAX) return fail(BLDFM_ELIMIT, "grid too large");
  if (nx) *nx = (int32_t)x;
  if (ny) *ny = (int32_t)y;
  return 0;
}

// Scalars of Eqs. (9), (11), (18), (19), (21), (31)-(36) of Kormann & Meixner (2001).
static void km01_scalars(const bldfm_km01& p, Km01Dev& D) {
  const double k = 0.4, zm = p.zm, L = p.mo_len;
  double phim, phic, psim, n;
  if (L < 0) {
    const double a = 1.0 - 16.0 * zm / L;
    phim = pow(a, -0.25);
    phic = pow(a, -0.5);
    const double x = pow(a, 0.25);
    psim = -2.0 * log(0.5 * (1.0 + x)) - log(0.5 * (1.0 + x * x)) + 2.0 * atan(x) - BLDFM_PI * 0.5;
    n = (1.0 - 24.0 * zm / L) / a;
  } else {
    phim = 1.0 + 5.0 * zm / L;
    phic = phim;
    psim = 5.0 * zm / L;
    n = 1.0 / (1.0 + 5.0 * zm / L);
  }
  const double m = p.ustar * phim / (k * p.ws);
  const double kappa = k * zm * p.ustar / (phic * pow(zm, n));
  const double U = p.ustar * (log(zm / p.z0) + psim) / (k * pow(zm, m));
  D.valid = U >= 0.0 ? 1 : 0;
  if (!D.valid) return;
  const double r = 2.0 + m - n, mu = (1.0 + m) / r, Xi = U * pow(zm, r) / (r * r * kappa);
  const double mr = m / r;
  const double A = U / (std::tgamma(1.0 / r) * p.sigma_v) * pow(kappa * r * r / U, mr);
  const double num = (1.0 / sqrt(2.0 * BLDFM_PI)) * pow(Xi, mu);
  D.scale = p.res * p.res * num * A;
  D.e1 = mr - 2.0 - mu;
  D.e2 = mr - 1.0;
  D.Xi = Xi;
  D.gA = std::tgamma(mu) * A;
}

int bldfm_km01_footprint(int device, const bldfm_km01* p, int32_t n, double* ffm, int32_t* status,
                         hipStream_t stream) {
  if (!p || n <= 0 || !ffm) return fail(BLDFM_EINVAL, "bad km01 arguments");
  std::vector<Km01Dev> D(n);
  int64_t off = 0, maxc = 0;
  for (int i = 0; i < n; ++i) {
    int32_t nx = 0, ny = 0;
    const int rc = bldfm_km01_grid(p + i, &nx, &ny);
    if (rc) return rc;
    Km01Dev& d = D[i];
    memset(&d, 0, sizeof(d));
    d.xmin = p[i].xmin; d.ymax = p[i].ymax; d.res = p[i].res; d.mx = p[i].mx; d.my = p[i].my;
    d.has_wd = p[i].has_wd ? 1 : 0;
    d.wd_rad = p[i].wd * (BLDFM_PI / 180.0);
    d.nx = nx; d.ny = ny; d.off = off;
    km01_scalars(p[i], d);
    if (status) status[i] = d.valid ? 0 : 1;
    off += (int64_t)nx * ny;
    maxc = std::max(maxc, (int64_t)nx * ny);
  }
  DeviceGuard guard(device);
  if (!guard.ok) return fail(BLDFM_EHIP, "hipSetDevice failed");
  void* dd = nullptr;
  hipError_t e = hipMalloc(&dd, sizeof(Km01Dev) * n);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(km01 records)");
  e = hipMemcpyAsync(dd, D.data(), sizeof(Km01Dev) * n, hipMemcpyHostToDevice, stream);
  int rc = 0;
  if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync(km01 records)");
  if (!rc) {
    const int lr = bldfm_launch_km01((const Km01Dev*)dd, n, maxc, ffm, stream);
    if (lr) rc = launch_fail(lr, "km01");
  }
  e = hipStreamSynchronize(stream);  // the records and their host copy must outlive the kernel
  if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  (void)hipFree(dd);
  return rc;
}

}  // extern "C"
