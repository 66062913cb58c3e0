// plan.cpp -- planner and C ABI of the gfx950 FFT engine (include/genfft_amd.h).
//
// A plan is a fixed list of kernel launches over three buffers (the caller's
// input and output and a plan-owned scratch), built once at creation:
//
//   C2C, n <= 16384 (c64) / 8192 (c128)  one batched row pass, one HBM round trip
//   C2C, larger n = N1 * N2                column pass (N1-point FFTs down N2 columns,
//                                          times W_n^(n2*k1)) into scratch, then a
//                                          transposing row pass (N2-point FFTs written
//                                          to k1 + N1*k2); the batch is processed in
//                                          chunks so the scratch ring stays small
//   R2C                                    C2C of n/2 on the reals viewed as complex,
//                                          then the unpack pass
//   C2R                                    pack pass, then an inverse C2C of n/2
//   DIT                                    the unpack pass alone
//   COLS                                   column tiles (two sub-passes when long)
//   C2C_2D                                 row pass, then column pass(es)
//   R2C_2D                                 row R2C on every row, column C2C on the
//                                          first width/2+1 columns, Hermitian fill
//
// Twiddles: every table is computed in long double with exact integer
// reduction of the angle and rounded once to the working precision; the
// four-step and real-FFT twiddles come from two-level FP64 tables multiplied
// in FP64 on the device, then rounded once.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/genfft_amd.h"
#include "gfft_batch.h"
#include "gfft_filter.h"
#include "gfft_misc.h"
#include "gfft_r2c.h"
#include "gfft_registry.h"

// Workspace of gfft_execute_noalloc: one static device array, instantiated
// on every GPU the library's code object is loaded on and held until the
// process ends.  Host code must never pass this name to a kernel: on the host
// it is only a stand-in the runtime resolves to each device's own copy
// (noalloc_scratch_ptr below).
static __device__ __align__(256) unsigned char noalloc_scratch[GFFT_NOALLOC_SCRATCH_BYTES];

namespace gfft {
namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(GFFT_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));        \
    } while (0)

bool is_pow2(int64_t n) { return n >= 1 && (n & (n - 1)) == 0; }
int ilog2(int64_t n)
{
    int l = 0;
    while (((int64_t)1 << l) < n) l++;
    return l;
}

// ---------------------------------------------------------------------------
// Twiddles: exp(-2*pi*i*m/n) in long double, m reduced mod n and folded into
// the first octant before cosl / sinl.
// ---------------------------------------------------------------------------
void twiddle_ld(int64_t m, int64_t n, long double *c, long double *s)
{
    static const long double PI = 3.141592653589793238462643383279502884197L;
    m %= n;
    if (m < 0) m += n;
    const int64_t a = 8 * m, q = a / n, r = a - q * n;
    const int64_t rr = (q & 1) ? (n - r) : r;
    const long double x = (2.0L * PI * (long double)rr) / (8.0L * (long double)n);
    const long double cx = cosl(x), sx = sinl(x);
    long double C, S;
    switch (q) {
    case 0: C = cx; S = sx; break;
    case 1: C = sx; S = cx; break;
    case 2: C = -sx; S = cx; break;
    case 3: C = -cx; S = sx; break;
    case 4: C = -cx; S = -sx; break;
    case 5: C = -sx; S = -cx; break;
    case 6: C = sx; S = -cx; break;
    default: C = cx; S = -sx; break;
    }
    *c = C;
    *s = -S;
}

struct DeviceTables {
    std::mutex mu;
    // (device, prec, L, radices) -> pass twiddles
    std::map<std::tuple<int, int, int, std::vector<int>>, void *> pass;
    // (device, N) -> two-level FP64 table {lo, hi}
    std::map<std::pair<int, int64_t>, std::pair<void *, void *>> two_level;
};

DeviceTables &tables()
{
    static DeviceTables *t = new DeviceTables();  // lives for the process
    return *t;
}

int upload(void **dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMalloc(dst, bytes));
    HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return GFFT_OK;
}

// Stockham pass twiddles of one tile shape: for each pass with NS > 1,
// entries [m*(R-1) + r-1] = W_{NS*R}^(r*m), m < NS, r = 1..R-1.
int get_pass_twiddles(int dev, const KernelEntry &k, const void **out, bool dry)
{
    if (dry) {
        *out = nullptr;
        return GFFT_OK;
    }
    auto &T = tables();
    std::lock_guard<std::mutex> lk(T.mu);
    std::vector<int> rad(k.rad, k.rad + k.nrad);
    auto key = std::make_tuple(dev, k.prec, k.L, rad);
    auto it = T.pass.find(key);
    if (it != T.pass.end()) {
        *out = it->second;
        return GFFT_OK;
    }
    std::vector<long double> re, im;
    int64_t ns = 1;
    for (int p = 0; p < k.nrad; p++) {
        const int R = k.rad[p];
        if (ns > 1)
            for (int64_t m = 0; m < ns; m++)
                for (int r = 1; r < R; r++) {
                    long double c, s;
                    twiddle_ld((int64_t)r * m, ns * R, &c, &s);
                    re.push_back(c);
                    im.push_back(s);
                }
        ns *= R;
    }
    if (re.empty()) {  // single-pass shapes read no twiddles; keep a valid pointer
        re.push_back(1);
        im.push_back(0);
    }
    void *d = nullptr;
    int st;
    if (k.prec) {
        std::vector<double> h(2 * re.size());
        for (size_t i = 0; i < re.size(); i++) h[2 * i] = (double)re[i], h[2 * i + 1] = (double)im[i];
        st = upload(&d, h.data(), h.size() * sizeof(double));
    } else {
        std::vector<float> h(2 * re.size());
        for (size_t i = 0; i < re.size(); i++) h[2 * i] = (float)re[i], h[2 * i + 1] = (float)im[i];
        st = upload(&d, h.data(), h.size() * sizeof(float));
    }
    if (st) return st;
    T.pass[key] = d;
    *out = d;
    return GFFT_OK;
}

// W_N^m for m <= N/2 as one FP64 table (N not a power of two): the same
// TwTable form with a mask of all ones and an empty high level, so
// hi[m >> lo_bits] = hi[0] = 1 exactly.
int get_direct(int dev, int64_t N, TwTable *t, bool dry)
{
    auto &T = tables();
    std::lock_guard<std::mutex> lk(T.mu);
    t->mask = ~0ull;
    t->lo_bits = 62;
    t->pad = 0;
    if (dry) {
        t->lo = t->hi = nullptr;
        return GFFT_OK;
    }
    auto key = std::make_pair(dev, -N);  // negative key: the direct tables
    auto it = T.two_level.find(key);
    if (it != T.two_level.end()) {
        t->lo = (const double *)it->second.first;
        t->hi = (const double *)it->second.second;
        return GFFT_OK;
    }
    const int64_t cnt = N / 2 + 1;
    std::vector<double> lo(2 * cnt), hi = {1.0, 0.0};
    for (int64_t m = 0; m < cnt; m++) {
        long double c, s;
        twiddle_ld(m, N, &c, &s);
        lo[2 * m] = (double)c;
        lo[2 * m + 1] = (double)s;
    }
    void *dlo = nullptr, *dhi = nullptr;
    int st = upload(&dlo, lo.data(), lo.size() * sizeof(double));
    if (st) return st;
    st = upload(&dhi, hi.data(), hi.size() * sizeof(double));
    if (st) return st;
    T.two_level[key] = {dlo, dhi};
    t->lo = (const double *)dlo;
    t->hi = (const double *)dhi;
    return GFFT_OK;
}

int get_two_level(int dev, int64_t N, TwTable *t, bool dry)
{
    auto &T = tables();
    std::lock_guard<std::mutex> lk(T.mu);
    const int lg = ilog2(N);
    const int lo_bits = (lg + 1) / 2;
    t->mask = (uint64_t)(N - 1);
    t->lo_bits = lo_bits;
    t->pad = 0;
    if (dry) {
        t->lo = t->hi = nullptr;
        return GFFT_OK;
    }
    auto key = std::make_pair(dev, N);
    auto it = T.two_level.find(key);
    if (it != T.two_level.end()) {
        t->lo = (const double *)it->second.first;
        t->hi = (const double *)it->second.second;
        return GFFT_OK;
    }
    const int64_t nlo = (int64_t)1 << lo_bits, nhi = std::max<int64_t>(1, N >> lo_bits);
    std::vector<double> lo(2 * nlo), hi(2 * nhi);
    for (int64_t m = 0; m < nlo; m++) {
        long double c, s;
        twiddle_ld(m, N, &c, &s);
        lo[2 * m] = (double)c;
        lo[2 * m + 1] = (double)s;
    }
    for (int64_t h = 0; h < nhi; h++) {
        long double c, s;
        twiddle_ld(h << lo_bits, N, &c, &s);
        hi[2 * h] = (double)c;
        hi[2 * h + 1] = (double)s;
    }
    void *dlo = nullptr, *dhi = nullptr;
    int st = upload(&dlo, lo.data(), lo.size() * sizeof(double));
    if (st) return st;
    st = upload(&dhi, hi.data(), hi.size() * sizeof(double));
    if (st) return st;
    T.two_level[key] = {dlo, dhi};
    t->lo = (const double *)dlo;
    t->hi = (const double *)dhi;
    return GFFT_OK;
}

// ---------------------------------------------------------------------------
// Kernel selection
// ---------------------------------------------------------------------------
int64_t env_int(const char *name, int64_t dflt);

// Per-plan kernel choices: the planner's defaults, an autotuned entry from the
// tuning table, or GFFT_* environment overrides (experiments).
struct PlanOpts {
    int n1_log2 = 0;       // four-step column length 2^n1_log2 (0: default split)
    int col_m = 0;         // preferred column-tile width (0: widest)
    int row_m = 0;         // preferred transposing-row tile width (0: widest)
    int pipe_mask = 3;     // pipelined kernels allowed: bit 0 column tiles, bit 1 transposing rows
    int row_variant = -1;  // registry index of the single-pass row kernel (-1: default)
    int r2c_fuse = 1;      // half-spectrum real FFTs: fused row-and-unpack pass where the shape allows
    int col_runs = -2;     // persistent column passes' tile order (TileArgs::runs; -2: by element size)
    int row_runs = 0;      // persistent transposing-row and single passes' tile order
    int dual = 0;          // four-step batches in chunks, column pass of chunk k beside row pass of k - 1
};

// Picks the tile kernel for (precision, L, patterns): the registry entry
// `variant` when it matches, else the widest tile (or the requested M) in its
// persistent pipelined form when one is compiled and allowed by pipe_mask.
const KernelEntry *tile_table(int prec, int store, int *cnt)
{
    if (store == 1) return tile_kernels_f16(cnt);
    if (store == 2) return tile_kernels_bf16(cnt);
    return prec ? tile_kernels_f64(cnt) : tile_kernels_f32(cnt);
}

const KernelEntry *find_tile(int prec, int L, int inp, int outp, int prefer_m = 0, int pipe_mask = 3,
                             int variant = -1, int store = 0)
{
    int cnt = 0;
    const KernelEntry *tab = tile_table(prec, store, &cnt);
    if (variant >= 0 && variant < cnt && tab[variant].L == L && tab[variant].inp == inp && tab[variant].outp == outp)
        return &tab[variant];
    const bool pipe = (inp == PAT_COL ? (pipe_mask & 1) : (pipe_mask & 2)) != 0;
    const KernelEntry *best = nullptr;
    auto better = [&](const KernelEntry &e) {
        if (!best) return true;
        if (prefer_m && (e.M == prefer_m) != (best->M == prefer_m)) return e.M == prefer_m;
        if (e.M != best->M) return !prefer_m && e.M > best->M;
        return pipe ? e.persistent > best->persistent : e.persistent < best->persistent;
    };
    for (int i = 0; i < cnt; i++) {
        const KernelEntry &e = tab[i];
        if (e.L != L || e.inp != inp || e.outp != outp || e.tune_only) continue;
        if (better(e)) best = &e;
    }
    return best;
}

// A batch whose default row tiles would number fewer than the CUs takes the
// narrowest registered row -> row tile of its length (tuning alternatives
// included, the one-wavefront and double-width forms aside) when that holds
// fewer rows than the default: more, shorter tiles side by side.  Measured on
// MI355X for complex64 1024-point rows: 16 rows 5.54 -> 5.14 us, 256 rows
// 5.84 -> 5.29 us (2 rows and 128 threads against 4 and 256).  -1 otherwise.
int narrow_row_variant(int prec, int L, int64_t batch, int num_cus)
{
    const KernelEntry *d = find_tile(prec, L, PAT_ROW, PAT_ROW);
    if (!d || batch / d->M >= num_cus) return -1;
    int cnt = 0, best = -1;
    const KernelEntry *tab = tile_table(prec, 0, &cnt);
    for (int i = 0; i < cnt; i++) {
        const KernelEntry &e = tab[i];
        if (e.L != L || e.inp != PAT_ROW || e.outp != PAT_ROW || e.wave || e.wide || e.persistent) continue;
        if (e.M < d->M && (best < 0 || e.M < tab[best].M)) best = i;
    }
    return best;
}

// Registry index of the row -> row entry of length L with `flag` set
// (KernelEntry::wave or ::wide), or -1.
int row_variant_of(int prec, int L, int KernelEntry::*flag)
{
    int cnt = 0;
    const KernelEntry *tab = tile_table(prec, 0, &cnt);
    for (int i = 0; i < cnt; i++)
        if (tab[i].*flag && tab[i].L == L && tab[i].inp == PAT_ROW && tab[i].outp == PAT_ROW) return i;
    return -1;
}

// The dual kernel pairing two pipelined tile kernels (role A, role B), if any.
const DualEntry *find_dual(int prec, const KernelEntry *a, const KernelEntry *b)
{
    if (!a || !b || !a->persistent || !b->persistent || a->store || b->store) return nullptr;
    int cnt = 0;
    const DualEntry *tab = prec ? dual_kernels_f64(&cnt) : dual_kernels_f32(&cnt);
    for (int i = 0; i < cnt; i++)
        if (tab[i].a_fn == a->fn && tab[i].b_fn == b->fn) return &tab[i];
    return nullptr;
}

// Narrowest tile (fewest sequences) registered for (L, inp, outp), tuning
// alternatives aside; 0 if none.
int min_m(int prec, int L, int inp, int outp, int store = 0)
{
    int cnt = 0, best = 0;
    const KernelEntry *tab = tile_table(prec, store, &cnt);
    for (int i = 0; i < cnt; i++)
        if (tab[i].L == L && tab[i].inp == inp && tab[i].outp == outp && !tab[i].tune_only)
            best = best ? std::min(best, tab[i].M) : tab[i].M;
    return best;
}

int max_len(int prec, int inp, int outp, int store = 0)
{
    int cnt = 0, best = 0;
    const KernelEntry *tab = tile_table(prec, store, &cnt);
    for (int i = 0; i < cnt; i++)
        if (tab[i].inp == inp && tab[i].outp == outp) best = std::max(best, tab[i].L);
    return best;
}

// Longest column tile whose rows are at least 64 bytes, the limit of the
// four-step column pass.  The 4096-point complex64 column tile (4 columns,
// 32-byte rows) serves column transforms of 4096 rows (the 2D 4096 x 4096
// column pass: 0.082 ms against 0.093 ms for two 64-point sub-passes), but
// a 4096 x 1024 four-step of 2^22 ran 1.165 ms against 1.115 ms for
// 2048 x 2048 on MI355X.
int max_fourstep_col(int prec, int store = 0)
{
    int cnt = 0, best = 0;
    const KernelEntry *tab = tile_table(prec, store, &cnt);
    const int cs = store ? 4 : (prec ? 16 : 8);
    for (int i = 0; i < cnt; i++)
        if (tab[i].inp == PAT_COL && tab[i].outp == PAT_COL && tab[i].M * cs >= 64) best = std::max(best, tab[i].L);
    return best;
}

int64_t env_int(const char *name, int64_t dflt)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    return strtoll(v, nullptr, 10);
}

// Four-step split of a length 2^lg: column pass 2^l1, transposing row pass 2^l2.
void four_step_split(int prec, int lg, int *l1, int *l2, int n1_override = 0)
{
    int b = lg / 2, a = lg - b;  // keep the column length <= the row length
    if (a > b) std::swap(a, b);
    // The transposing row pass runs at copy speed at length 1024 (pipelined
    // kernel) while column passes are fastest short, so keep the row length
    // at 1024 while the column length fits a tile.  Measured on MI355X for
    // 2 GiB complex64 batches: 2^16 0.79 ms, 2^18 0.81 ms, 2^20 0.92 ms; for
    // 1 GiB complex128 batches: 2^14 1.04 -> 0.81 ms, 2^18 1.16 -> 0.86 ms
    // against the even split.
    if (lg - 10 >= 1 && ((int64_t)1 << (lg - 10)) <= max_fourstep_col(prec)) {
        b = 10;
        a = lg - 10;
    }
    if (n1_override > 0 && n1_override < lg) {
        a = n1_override;
        b = lg - a;
    }
    *l1 = a;
    *l2 = b;
}

// ---------------------------------------------------------------------------
// Tuning table: (precision, C2C length) -> kernel choices found by
// gfft_autotune, optionally loaded from the JSON file named by
// GFFT_TUNING_FILE at the first plan creation.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::pair<int, int64_t>, PlanOpts> g_tune;
bool g_tune_loaded = false;

int tuning_load_locked(const char *path);

PlanOpts resolve_opts(int prec, int64_t n, const PlanOpts *forced)
{
    if (forced) return *forced;
    PlanOpts o;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (!g_tune_loaded) {
            g_tune_loaded = true;
            if (const char *f = getenv("GFFT_TUNING_FILE"))
                if (*f) (void)tuning_load_locked(f);
        }
        auto it = g_tune.find({prec, n});
        if (n > 0 && it != g_tune.end()) o = it->second;
    }
    // environment overrides (experiments)
    if (getenv("GFFT_N1_LOG2")) o.n1_log2 = (int)env_int("GFFT_N1_LOG2", 0);
    if (getenv("GFFT_COL_M")) o.col_m = (int)env_int("GFFT_COL_M", 0);
    if (getenv("GFFT_ROW_M")) o.row_m = (int)env_int("GFFT_ROW_M", 0);
    if (getenv("GFFT_PIPE")) o.pipe_mask = (int)env_int("GFFT_PIPE", 3);
    if (getenv("GFFT_R2C_FUSE")) o.r2c_fuse = (int)env_int("GFFT_R2C_FUSE", 1);
    if (getenv("GFFT_COL_RUNS")) o.col_runs = (int)env_int("GFFT_COL_RUNS", -2);
    if (getenv("GFFT_ROW_RUNS")) o.row_runs = (int)env_int("GFFT_ROW_RUNS", 0);
    if (getenv("GFFT_DUAL")) o.dual = (int)env_int("GFFT_DUAL", 0);
    return o;
}

// ---------------------------------------------------------------------------
// Plan representation
// ---------------------------------------------------------------------------
enum BufId { B_IN = 0, B_OUT = 1, B_SCR = 2 };
enum LaunchType { LT_TILE, LT_UNPACK, LT_PACK, LT_HFILL, LT_COPY, LT_MIX, LT_MIX_RESET, LT_R2C, LT_C2R, LT_RROWS,
                  LT_REPACK, LT_DUAL };

// bytes of one complex element in memory
size_t csize_of(int prec, int store) { return store ? 4 : (prec ? 16 : 8); }

struct Launch {
    LaunchType type;
    const KernelEntry *k = nullptr;
    TileArgs ta{};
    UnpackArgs ua{};
    HFillArgs ha{};
    CopyArgs ca{};
    MixArgs mix{};
    R2cArgs ra{};
    C2rArgs cr{};
    RealRowsArgs rr{};
    RepackArgs rp{};
    int rr_mode = RR_R2C;
    int in_buf = B_IN, out_buf = B_OUT;
    int64_t in_off = 0, out_off = 0;  // bytes
    bool vin_ok = false, vout_ok = false;
    bool first = false, last = false;  // ends of a complex chain (inverse swaps)
    int64_t grid = 0;
    // LT_DUAL: a second tile pass in the same launch (role B; role A is the
    // tile pass described by the fields above), run by the workgroups from
    // `split` on
    const DualEntry *dual = nullptr;
    const KernelEntry *kb = nullptr;
    TileArgs tb{};
    int in_buf_b = B_IN, out_buf_b = B_OUT;
    int64_t in_off_b = 0, out_off_b = 0;
    bool vin_ok_b = false, vout_ok_b = false, first_b = false, last_b = false;
    int64_t split = 0;
    bool reads(int buf) const { return in_buf == buf || (type == LT_DUAL && in_buf_b == buf); }
    bool writes(int buf) const { return out_buf == buf || (type == LT_DUAL && out_buf_b == buf); }
};

}  // namespace
}  // namespace gfft

// Orders the work of one plan across streams.  A plan's workspace (scratch,
// partial sums) is shared by all its calls, so a call on a new stream first
// waits for the plan's previous work; calls on one stream are ordered by the
// stream itself.  (genFFT's plans are reentrant across threads for distinct
// buffers, FFTLevel.h:76-92; this keeps the shared-workspace plans so.)
// The event only orders device work (hipStreamWaitEvent) and is never waited
// on by the host, so it is recorded without the system-scope fence: with the
// fence, each record wrote back and invalidated the caches, and the next
// kernel on the stream started ~5.6 us after the plan's last one ended (MI355X
// kernel traces, every call of a plan with a workspace).
struct StreamOrder {
    std::mutex mu;
    hipEvent_t done = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
    ~StreamOrder()
    {
        if (done) (void)hipEventDestroy(done);
    }
    // call with mu held, before the plan's launches on s
    hipError_t before(hipStream_t s)
    {
        if (!done) {
            hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming | hipEventDisableSystemFence);
            if (e != hipSuccess) return e;
        }
        if (used && last != s) return hipStreamWaitEvent(s, done, 0);
        return hipSuccess;
    }
    // call with mu held, after the plan's launches on s
    hipError_t after(hipStream_t s)
    {
        mark(s);
        return hipEventRecord(done, s);
    }
    // instead of after(), when the plan's last launch on s carried `done` as
    // its stop event (hipExtLaunchKernel): the event then completes with that
    // kernel, and no marker packet sits between it and the stream's next
    // kernel (one cost ~2 us per call on MI355X)
    void mark(hipStream_t s)
    {
        last = s;
        used = true;
    }
    // the event to attach to the last launch on s; null while s is being
    // captured into a graph (the capture then records it as a node, after())
    hipEvent_t stop_event(hipStream_t s) const
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
        return done;
    }
};

struct gfft_plan_s {
    gfft_desc d{};
    int prec = 0;   // arithmetic: 0 FP32, 1 FP64
    int store = 0;  // memory format: 0 complex of the arithmetic type, 1 FP16 pairs, 2 BF16 pairs
    int device = 0;
    // per-transform scale (16-bit storage): transform b starts at element
    // b * scale_div of the input; the auto-scale pass reads its rows x width
    // elements (row pitch amax_pitch) and bounds every stored value by
    // amax_gain * max(|re| + |im|)
    int64_t scale_div = 1, amax_rows = 1, amax_width = 1, amax_pitch = 1, amax_gain = 1;
    std::vector<gfft::Launch> launches;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int dir_mask = 3;  // bit0 forward allowed, bit1 inverse allowed
    int xcd_remap = 1;
    int num_cus = 256;
    size_t in_bytes = 0, out_bytes = 0;
    bool out_has_gaps = false;
    // Set when some launch may write part of the output before all of the
    // input is read: a launch that reads the input and writes the output
    // (its workgroups store while others have yet to load), or an output
    // write ahead of a later input read (a batch run in chunks).  Overlapping
    // buffers are then staged (needs_stage) unless every transform is in
    // place: in == out with equal byte distances in_step / out_step (batch
    // members, or rows) and in_step2 / out_step2 (2D images).
    bool overlap_hazard = false;
    int64_t in_step = 0, out_step = 0, in_step2 = 0, out_step2 = 0;
    void *stage = nullptr;  // copy of an input that overlaps the output (allocated on first use)
    std::mutex host_mu;
    void *hin = nullptr, *hout = nullptr;
    // small host calls: pinned copies of both sides and a stream of their own
    void *pin_in = nullptr, *pin_out = nullptr;
    hipStream_t host_stream = nullptr;
    StreamOrder order;  // used when the plan has a workspace
    void *perm = nullptr;  // natural-order copy of a scrambled input (gfft_execute_scrambled)
};

namespace gfft {
namespace {

// The 16-byte paths of the complex64 tile kernels address an access as a
// uniform base plus one 32-bit per-thread byte offset, so every in-tile
// offset must fit 31 bits (element n of sequence s lies s * inner + n (ROW)
// or n * elem + s (COL) from the tile's first element).
bool tile_offsets_fit(int pat, int64_t L, int64_t M, int64_t inner, int64_t elem, size_t csize)
{
    if (inner < 0 || elem < 0) return false;
    const int64_t last = pat == PAT_ROW ? (M - 1) * inner + (L - 1) : (L - 1) * elem + (M - 1);
    return last * (int64_t)csize < ((int64_t)1 << 31);
}

struct Builder {
    gfft_plan_s *p;
    int prec;
    size_t csize;  // bytes of one complex element
    bool dry;      // plan without a device: no table uploads
    const PlanOpts *forced = nullptr;  // autotuner candidate (overrides table and environment)
    PlanOpts opts = resolve_opts(prec, 0, forced);
    int64_t scratch_elems = 0;  // complex elements of scratch needed
    int store = 0;              // 16-bit storage format (tile_kernels_f16 / _bf16), default shapes only

    void need_scratch(int64_t elems) { scratch_elems = std::max(scratch_elems, elems); }

    // Largest batch chunk a four-step plan runs through its workspace: 1 GiB
    // (GFFT_CHUNK_BYTES overrides), or the static workspace of
    // gfft_execute_noalloc for plans created without a workspace.
    int64_t chunk_bytes() const
    {
        int64_t c = env_int("GFFT_CHUNK_BYTES", (int64_t)1 << 30);
        if (p->d.flags & GFFT_FLAG_NO_WORKSPACE) c = std::min<int64_t>(c, GFFT_NOALLOC_SCRATCH_BYTES);
        return c;
    }

    // one tile pass
    int tile(int L, int inp, int outp, int prefer_m, int in_buf, int64_t in_off_el, int out_buf,
             int64_t out_off_el, int64_t n_inner, int64_t n_outer1, int64_t n_outer2,
             int64_t in_outer2, int64_t in_outer, int64_t in_inner, int64_t in_elem,
             int64_t out_outer2, int64_t out_outer, int64_t out_inner, int64_t out_elem,
             int tw_mode, int64_t tw_n, bool first, bool last, int variant = -1)
    {
        const KernelEntry *k = find_tile(prec, L, inp, outp, prefer_m, opts.pipe_mask, variant, store);
        if (!k) return fail(GFFT_ERR_UNSUPPORTED, "no tile kernel for L=%d pattern %d->%d", L, inp, outp);
        Launch l;
        l.type = LT_TILE;
        l.k = k;
        TileArgs &a = l.ta;
        a.in_outer2 = in_outer2;
        a.in_outer = in_outer;
        a.in_inner = in_inner;
        a.in_elem = in_elem;
        a.out_outer2 = out_outer2;
        a.out_outer = out_outer;
        a.out_inner = out_inner;
        a.out_elem = out_elem;
        const int64_t tiles_inner = (n_inner + k->M - 1) / k->M;
        const int64_t ntiles = tiles_inner * n_outer1 * n_outer2;
        if (ntiles > INT32_MAX || n_inner > INT32_MAX)
            return fail(GFFT_ERR_UNSUPPORTED, "too many tiles (%lld)", (long long)ntiles);
        a.n_inner = (int32_t)n_inner;
        a.tiles_inner = (int32_t)tiles_inner;
        a.n_outer1 = (int32_t)n_outer1;
        a.n_tiles = (int32_t)ntiles;
        a.tw_mode = tw_mode;
        if (tw_mode != TW_NONE) {
            TwTable t;
            int st = get_two_level(p->device, tw_n, &t, dry);
            if (st) return st;
            a.tw_lo = t.lo;
            a.tw_hi = t.hi;
            a.tw_mask = (uint32_t)t.mask;
            a.tw_lo_bits = t.lo_bits;
        }
        if (inp == PAT_ROW ? in_elem != 1 : in_inner != 1)
            return fail(GFFT_ERR_UNSUPPORTED, "internal: input pattern does not match strides");
        if (outp == PAT_ROW ? out_elem != 1 : out_inner != 1)
            return fail(GFFT_ERR_UNSUPPORTED, "internal: output pattern does not match strides");
        int st = get_pass_twiddles(p->device, *k, &a.pass_tw, dry);
        if (st) return st;
        if (k->gen_tw) {
            TwTable g;
            if ((st = get_two_level(p->device, k->L, &g, dry))) return st;
            a.gen_lo = g.lo;
            a.gen_hi = g.hi;
        }
        // 16-byte vectors need even element offsets everywhere in the tile
        auto even = [](int64_t v) { return (v & 1) == 0; };
        if (prec == 0 && store == 0) {
            l.vin_ok = even(in_outer2) && even(in_outer) && even(in_off_el) &&
                       (inp == PAT_ROW ? even(in_inner) : even(in_elem)) &&
                       tile_offsets_fit(inp, k->L, k->M, in_inner, in_elem, csize);
            l.vout_ok = even(out_outer2) && even(out_outer) && even(out_off_el) &&
                        (outp == PAT_ROW ? even(out_inner) : even(out_elem)) &&
                        tile_offsets_fit(outp, k->L, k->M, out_inner, out_elem, csize);
        } else if (store) {
            // 16-bit pairs: a 16-byte vector is four elements
            auto q4 = [](int64_t v) { return (v & 3) == 0; };
            l.vin_ok = q4(in_outer2) && q4(in_outer) && q4(in_off_el) && (inp == PAT_ROW ? q4(in_inner) : q4(in_elem));
            l.vout_ok = q4(out_outer2) && q4(out_outer) && q4(out_off_el) &&
                        (outp == PAT_ROW ? q4(out_inner) : q4(out_elem));
        }
        l.in_buf = in_buf;
        l.out_buf = out_buf;
        l.in_off = in_off_el * (int64_t)csize;
        l.out_off = out_off_el * (int64_t)csize;
        l.first = first;
        l.last = last;
        l.grid = ntiles;
        if (k->persistent) {
            // one resident wave of workgroups looping over the tiles
            int per_cu = 1;
            if (!dry && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k->fn, k->NT, 0) !=
                            hipSuccess)
                per_cu = 1;
            l.grid = std::min<int64_t>(ntiles, (int64_t)std::max(1, per_cu) * std::max(1, p->num_cus));
            // contiguous runs of tiles per workgroup for the column passes of
            // 8- and 16-byte elements (measured faster); interleaved otherwise
            int runs = inp == PAT_COL ? (opts.col_runs == -2 ? (csize >= 8 ? -1 : 0) : opts.col_runs)
                                      : opts.row_runs;
            if (runs < -1 || runs == 1 || (runs > 1 && (runs & (runs - 1)))) runs = 0;  // powers of two
            a.runs = runs;
            // the kernel steps a tile index by up to grid * run length past the last tile
            if (ntiles + l.grid * std::max(1, runs) > INT32_MAX)
                return fail(GFFT_ERR_UNSUPPORTED, "too many tiles (%lld)", (long long)ntiles);
        }
        p->launches.push_back(l);
        return GFFT_OK;
    }

    // Batched 1D complex transforms of length n; `batch` transforms at
    // distances in_dist / out_dist (complex elements).
    int c2c_1d(int64_t n, int64_t batch, int in_buf, int64_t in_off, int64_t in_dist, int out_buf,
               int64_t out_off, int64_t out_dist, bool first, bool last)
    {
        if (n == 1) {
            Launch l;
            l.type = LT_COPY;
            l.ca.batch = batch;
            l.ca.in_dist = in_dist;
            l.ca.out_dist = out_dist;
            l.in_buf = in_buf;
            l.out_buf = out_buf;
            l.in_off = in_off * (int64_t)csize;
            l.out_off = out_off * (int64_t)csize;
            l.first = first;
            l.last = last;
            p->launches.push_back(l);
            return GFFT_OK;
        }
        const int maxrow = max_len(prec, PAT_ROW, PAT_ROW, store);
        const PlanOpts saved = opts;
        opts = store ? PlanOpts{} : resolve_opts(prec, n, forced);
        struct Restore {
            PlanOpts &dst;
            PlanOpts v;
            ~Restore() { dst = v; }
        } restore{opts, saved};
        if (n <= maxrow) {
            // Without a tuned choice: a single row takes the one-wavefront
            // kernel where one is compiled for this length (no barrier; c64
            // n = 1024 on MI355X: 4.2 against 5.4 us for the LDS tile, which
            // wins from 16 rows on), a batch of >= 512 MiB the double-width
            // tile where one is registered (KernelEntry::wide).
            int variant = opts.row_variant;
            if (variant < 0 && !store) {
                if (batch == 1) variant = row_variant_of(prec, (int)n, &KernelEntry::wave);
                else if (n * batch * (int64_t)csize >= ((int64_t)512 << 20))
                    variant = row_variant_of(prec, (int)n, &KernelEntry::wide);
                else variant = narrow_row_variant(prec, (int)n, batch, p->num_cus);
            }
            return tile((int)n, PAT_ROW, PAT_ROW, 0, in_buf, in_off, out_buf, out_off, batch, 1, 1, 0, 0,
                        in_dist, 1, 0, 0, out_dist, 1, TW_NONE, 0, first, last, variant);
        }
        // four-step: n = N1 (column length) * N2 (row length)
        const int lg = ilog2(n);
        int l1, l2;
        four_step_split(prec, lg, &l1, &l2, opts.n1_log2);
        const int64_t N1 = (int64_t)1 << l1, N2 = (int64_t)1 << l2;
        if (N1 > max_len(prec, PAT_COL, PAT_COL, store) || N2 > max_len(prec, PAT_ROW, PAT_COL, store))
            return fail(GFFT_ERR_UNSUPPORTED, "length %lld exceeds the two-pass limit", (long long)n);
        if (opts.dual && !store) {
            const int st = four_step_dual(n, N1, N2, batch, in_buf, in_off, in_dist, out_buf, out_off, out_dist, first,
                                          last);
            if (st != kNotApplicable) return st;
        }
        // Run one after the other, the two passes are each fastest as one
        // launch over the whole batch: chunks only bound the scratch size
        // (measured on MI355X, c64 2^20 x 128: sequential chunks through a
        // small ring gain nothing, launch tails outweigh any Infinity-Cache reuse).
        int64_t chunk = std::max<int64_t>(1, chunk_bytes() / (n * (int64_t)csize));
        chunk = std::min(chunk, batch);
        need_scratch(chunk * n);
        // Without a tuned choice, a pass whose default tiles would leave CUs
        // idle (small batches) takes the narrowest registered tile: more,
        // shorter tiles run side by side.
        auto narrow = [&](int L, int inp, int outp, int64_t seqs, int tuned) {
            if (tuned || store) return tuned;
            const KernelEntry *k = find_tile(prec, L, inp, outp, 0, opts.pipe_mask, -1, store);
            const int m = min_m(prec, L, inp, outp, store);
            return k && m < k->M && seqs / k->M < p->num_cus ? m : 0;
        };
        const int pm1 = narrow((int)N1, PAT_COL, PAT_COL, chunk * N2, opts.col_m);
        const int pm2 = narrow((int)N2, PAT_ROW, PAT_COL, chunk * N1, opts.row_m);
        for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
            const int64_t cb = std::min(chunk, batch - b0);
            int st = tile((int)N1, PAT_COL, PAT_COL, pm1, in_buf, in_off + b0 * in_dist, B_SCR, 0, N2, cb, 1, 0,
                          in_dist, 1, N2, 0, n, 1, N2, TW_INNER, n, first, false);
            if (st) return st;
            st = tile((int)N2, PAT_ROW, PAT_COL, pm2, B_SCR, 0, out_buf, out_off + b0 * out_dist, N1, cb, 1, 0,
                      n, N2, 1, 0, out_dist, 1, N1, TW_NONE, 0, false, last);
            if (st) return st;
        }
        return GFFT_OK;
    }

    static constexpr int kNotApplicable = -1;

    // The overlapped four-step: the batch in chunks of dc transforms through a
    // ring of two scratch slots; launch 0 runs the column pass of chunk 0,
    // launch k (1 <= k < chunks) pairs the column pass of chunk k with the
    // transposing row pass of chunk k - 1 (dual_pipe_kernel, the workgroups
    // split between the two), and the last launch runs the row pass of the
    // last chunk.  Launches on one stream run in order, so slot k & 1 is
    // written by launch k and read by launch k + 1 only after every write of
    // launch k: no wait inside a kernel.  Returns kNotApplicable when no dual
    // kernel pairs the two passes of this split, or the batch is one chunk.
    int four_step_dual(int64_t n, int64_t N1, int64_t N2, int64_t batch, int in_buf, int64_t in_off, int64_t in_dist,
                       int out_buf, int64_t out_off, int64_t out_dist, bool first, bool last)
    {
        const int64_t bytes = std::min<int64_t>(env_int("GFFT_DUAL_CHUNK_BYTES", (int64_t)64 << 20), chunk_bytes() / 2);
        const int64_t dc = std::max<int64_t>(1, bytes / (n * (int64_t)csize));
        if (dc >= batch) return kNotApplicable;
        const int pm1 = opts.col_m, pm2 = opts.row_m;
        const KernelEntry *kc = find_tile(prec, (int)N1, PAT_COL, PAT_COL, pm1, opts.pipe_mask, -1, store);
        const KernelEntry *kr = find_tile(prec, (int)N2, PAT_ROW, PAT_COL, pm2, opts.pipe_mask, -1, store);
        const DualEntry *d = find_dual(prec, kc, kr);
        if (!d) return kNotApplicable;
        const int64_t chunks = (batch + dc - 1) / dc;
        need_scratch(2 * dc * n);
        auto take = [&](Launch &l) {
            l = p->launches.back();
            p->launches.pop_back();
        };
        auto col = [&](int64_t c, Launch &l) {
            const int64_t cb = std::min(dc, batch - c * dc);
            const int st = tile((int)N1, PAT_COL, PAT_COL, pm1, in_buf, in_off + c * dc * in_dist, B_SCR,
                                (c & 1) * dc * n, N2, cb, 1, 0, in_dist, 1, N2, 0, n, 1, N2, TW_INNER, n, first, false);
            if (!st) take(l);
            return st;
        };
        auto row = [&](int64_t c, Launch &l) {
            const int64_t cb = std::min(dc, batch - c * dc);
            const int st = tile((int)N2, PAT_ROW, PAT_COL, pm2, B_SCR, (c & 1) * dc * n, out_buf,
                                out_off + c * dc * out_dist, N1, cb, 1, 0, n, N2, 1, 0, out_dist, 1, N1, TW_NONE, 0,
                                false, last);
            if (!st) take(l);
            return st;
        };
        int per_cu = 1;
        if (!dry && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)d->fn, d->NT, 0) != hipSuccess)
            per_cu = 1;
        const int64_t total = (int64_t)std::max(1, per_cu) * std::max(1, p->num_cus);
        // workgroups of the column role, in 256ths of the grid (measured with
        // copies in the two passes' access patterns: an even split is best)
        const int64_t share = std::min<int64_t>(255, std::max<int64_t>(1, env_int("GFFT_DUAL_SPLIT", 128)));
        const int64_t ga0 = std::max<int64_t>(1, std::min(total - 1, total * share / 256));
        Launch l, a, b;
        int st = col(0, l);
        if (st) return st;
        p->launches.push_back(l);
        for (int64_t c = 1; c < chunks; c++) {
            if ((st = col(c, a)) || (st = row(c - 1, b))) return st;
            a.type = LT_DUAL;
            a.dual = d;
            a.kb = b.k;
            a.tb = b.ta;
            a.in_buf_b = b.in_buf;
            a.out_buf_b = b.out_buf;
            a.in_off_b = b.in_off;
            a.out_off_b = b.out_off;
            a.vin_ok_b = b.vin_ok;
            a.vout_ok_b = b.vout_ok;
            a.first_b = b.first;
            a.last_b = b.last;
            a.split = std::min<int64_t>(ga0, a.ta.n_tiles);
            a.grid = a.split + std::min<int64_t>(total - ga0, b.ta.n_tiles);
            p->launches.push_back(a);
        }
        if ((st = row(chunks - 1, l))) return st;
        p->launches.push_back(l);
        return GFFT_OK;
    }

    // Column FFTs of length `rows` down `cols` columns, `batch` matrices.
    // in/out pitch (row stride) and batch distance in complex elements.
    int columns(int64_t rows, int64_t cols, int64_t batch, int in_buf, int64_t in_pitch, int64_t in_dist,
                int out_buf, int64_t out_pitch, int64_t out_dist, bool first, bool last)
    {
        if (rows == 1) {
            if (in_buf == out_buf && in_pitch == out_pitch && in_dist == out_dist) return GFFT_OK;
            // the length-1 transform is the identity: one strided copy of the
            // `cols` elements of every matrix, moved as 4- or 8-byte words
            const int32_t word = csize == 4 ? 4 : (int32_t)(csize / 2);
            const int64_t per = (int64_t)csize / word;
            Launch l;
            l.type = LT_REPACK;
            l.rp = RepackArgs{nullptr, nullptr, cols * per, 1, batch, 0, in_dist * per, 0, out_dist * per, word, 0};
            l.in_buf = in_buf;
            l.out_buf = out_buf;
            p->launches.push_back(l);
            return GFFT_OK;
        }
        const int maxcol = max_len(prec, PAT_COL, PAT_COL, store);
        if (rows <= maxcol)
            return tile((int)rows, PAT_COL, PAT_COL, 0, in_buf, 0, out_buf, 0, cols, batch, 1, 0, in_dist, 1,
                        in_pitch, 0, out_dist, 1, out_pitch, TW_NONE, 0, first, last);
        // rows = A * B, sub-pass 1 over n2 (length B, rows n1 + A*n2) with
        // twiddle W_rows^(n1*k2) into scratch, sub-pass 2 over n1 (length A)
        // writing row k2 + B*k1.
        const int lg = ilog2(rows);
        const int lb = lg / 2, la = lg - lb;
        const int64_t A = (int64_t)1 << la, B = (int64_t)1 << lb;
        if (A > maxcol) return fail(GFFT_ERR_UNSUPPORTED, "column length %lld too long", (long long)rows);
        const int64_t spitch = cols, sdist = rows * cols;
        need_scratch(sdist * batch);
        int st = tile((int)B, PAT_COL, PAT_COL, 0, in_buf, 0, B_SCR, 0, cols, A, batch, in_dist, in_pitch, 1,
                      A * in_pitch, sdist, spitch, 1, A * spitch, TW_OUTER, rows, first, false);
        if (st) return st;
        return tile((int)A, PAT_COL, PAT_COL, 0, B_SCR, 0, out_buf, 0, cols, B, batch, sdist, A * spitch, 1,
                    spitch, out_dist, out_pitch, 1, B * out_pitch, TW_NONE, 0, false, last);
    }

    int unpack(int64_t n, int64_t batch, int in_buf, int64_t in_dist, int out_buf, int64_t out_dist, bool half,
               bool pack)
    {
        Launch l;
        l.type = pack ? LT_PACK : LT_UNPACK;
        l.ua.n = n;
        l.ua.batch = batch;
        l.ua.in_dist = in_dist;
        l.ua.out_dist = out_dist;
        l.ua.half = half ? 1 : 0;
        if (n >= 2) {
            int st = is_pow2(n) ? get_two_level(p->device, n, &l.ua.tw, dry) : get_direct(p->device, n, &l.ua.tw, dry);
            if (st) return st;
        }
        l.in_buf = in_buf;
        l.out_buf = out_buf;
        p->launches.push_back(l);
        return GFFT_OK;
    }

    // Whether the half-spectrum real FFT of length n runs as a column pass
    // plus the fused row-and-unpack pass: n/2 = n1 * R with the n1-point
    // column pass in the table and whole workgroups of row pairs.
    bool r2c_fusable(int64_t n) const
    {
        const KernelEntry *fe = r2c_fused_entry(prec);
        const int64_t L = n / 2, R = fe->L;
        if (store || !is_pow2(n) || L <= max_len(prec, PAT_ROW, PAT_ROW) || L % R) return false;
        const int64_t n1 = L / R;
        return n1 >= fe->M && n1 % fe->M == 0 && n1 <= max_fourstep_col(prec);
    }

    // Half-spectrum real FFT of length n (n/2 = n1 * R): the n1-point
    // column pass of the n/2-point four-step over the whole batch into
    // scratch, then one launch that runs the R-point row FFTs and writes bins
    // 0 .. n/2 of the output.  Never chunked: every input element is read
    // before the first bin is written, so in == out is safe whatever the
    // distances (with the dense layout a transform's bins overlap the next
    // transform's input).
    int r2c_fused(int64_t n, int64_t batch, int64_t in_dist, int64_t out_dist)
    {
        const KernelEntry *fe = r2c_fused_entry(prec);
        const int64_t L = n / 2, R = fe->L, n1 = L / R;
        need_scratch(batch * L);
        int st = tile((int)n1, PAT_COL, PAT_COL, 0, B_IN, 0, B_SCR, 0, R, batch, 1, 0, in_dist / 2, 1, R, 0, L, 1, R,
                      TW_INNER, L, true, false);
        if (st) return st;
        Launch l;
        l.type = LT_R2C;
        l.k = fe;
        R2cArgs &r = l.ra;
        r.in_dist = L;
        r.out_dist = out_dist;
        r.n1 = (int32_t)n1;
        r.tiles_per = (int32_t)(n1 / fe->M);
        if ((int64_t)r.tiles_per * batch > INT32_MAX) return fail(GFFT_ERR_UNSUPPORTED, "too many tiles");
        r.n_tiles = (int32_t)(r.tiles_per * batch);
        if ((st = get_pass_twiddles(p->device, *fe, &r.pass_tw, dry))) return st;
        if ((st = get_two_level(p->device, n, &r.tw, dry))) return st;
        l.in_buf = B_SCR;
        l.out_buf = B_OUT;
        l.vin_ok = prec == 0;  // rows start at even element offsets of scratch
        l.last = true;
        l.grid = r.n_tiles;
        p->launches.push_back(l);
        return GFFT_OK;
    }

    // Whether a real transform of length n runs in one row-tile pass (n/2 a
    // single-pass row length from 2 on): unpack or pack fused in.
    bool rows_fusable(int64_t n) const
    {
        return !store && opts.r2c_fuse && is_pow2(n) && n / 2 >= 2 && real_rows_entry(prec, (int)(n / 2));
    }

    // `rows` real transforms of length n in one launch: mode RR_R2C reads
    // rows of n reals (in_row apart, complex units) and writes bins (out_row
    // apart; n/2 + 1 of them with `half`, else n); RR_C2R reads n/2 + 1 bins
    // per row and writes n reals (out_row apart, complex units).
    int real_rows(int mode, int64_t n, int64_t rows, int in_buf, int64_t in_off_el, int64_t in_row, int out_buf,
                  int64_t out_off_el, int64_t out_row, bool half)
    {
        const KernelEntry *fe = real_rows_entry(prec, (int)(n / 2));
        Launch l;
        l.type = LT_RROWS;
        l.k = fe;
        l.rr_mode = mode;
        TileArgs &a = l.rr.t;
        const int64_t tiles = (rows + fe->M - 1) / fe->M;
        if (tiles > INT32_MAX || rows > INT32_MAX) return fail(GFFT_ERR_UNSUPPORTED, "too many rows");
        a.in_inner = in_row;
        a.in_elem = 1;
        a.out_inner = out_row;
        a.out_elem = 1;
        a.n_inner = (int32_t)rows;
        a.tiles_inner = (int32_t)tiles;
        a.n_outer1 = 1;
        a.n_tiles = (int32_t)tiles;
        a.tw_mode = TW_NONE;
        a.swap_out = mode == RR_C2R;
        l.rr.half = half ? 1 : 0;
        int st;
        if ((st = get_pass_twiddles(p->device, *fe, &a.pass_tw, dry))) return st;
        if ((st = get_two_level(p->device, n, &l.rr.tw, dry))) return st;
        auto even = [](int64_t v) { return (v & 1) == 0; };
        if (prec == 0) {
            l.vin_ok = mode == RR_R2C && even(in_row) && even(in_off_el) &&
                       tile_offsets_fit(PAT_ROW, fe->L, fe->M, in_row, 1, csize);
            l.vout_ok = mode == RR_C2R && even(out_row) && even(out_off_el) &&
                        tile_offsets_fit(PAT_ROW, fe->L, fe->M, out_row, 1, csize);
        }
        l.in_buf = in_buf;
        l.out_buf = out_buf;
        l.in_off = in_off_el * (int64_t)csize;
        l.out_off = out_off_el * (int64_t)csize;
        l.grid = tiles;
        p->launches.push_back(l);
        return GFFT_OK;
    }

    // Whether the c2r of length n runs as the fused pack-and-column pass plus
    // the transposing row pass of the n/2-point inverse four-step.
    bool c2r_fusable(int64_t n) const
    {
        const int64_t L = n / 2;
        if (store || !is_pow2(n) || L <= max_len(prec, PAT_ROW, PAT_ROW)) return false;
        int l1, l2;
        four_step_split(prec, ilog2(L), &l1, &l2);
        const KernelEntry *fe = c2r_fused_entry(prec, 1 << l1);
        return fe && (1 << l2) % fe->M == 0 && (1 << l2) <= max_len(prec, PAT_ROW, PAT_COL);
    }

    // c2r of length n (n/2 = n1 * n2): the pack fused into the n1-point
    // column pass over the whole batch into scratch (every bin is read before
    // a real is written, so in == out is safe), then the n2-point transposing
    // row pass writes the reals.
    int c2r_fused(int64_t n, int64_t batch, int64_t in_dist, int64_t out_dist)
    {
        const int64_t L = n / 2;
        int l1, l2;
        four_step_split(prec, ilog2(L), &l1, &l2);
        const int64_t N1 = (int64_t)1 << l1, N2 = (int64_t)1 << l2;
        const KernelEntry *fe = c2r_fused_entry(prec, (int)N1);
        need_scratch(batch * L);
        Launch l;
        l.type = LT_C2R;
        l.k = fe;
        C2rArgs &c = l.cr;
        c.in_dist = in_dist;
        c.n2 = (int32_t)N2;
        c.tiles_per = (int32_t)(N2 / fe->M);
        if ((int64_t)c.tiles_per * batch > INT32_MAX) return fail(GFFT_ERR_UNSUPPORTED, "too many tiles");
        c.n_tiles = (int32_t)(c.tiles_per * batch);
        int st;
        if ((st = get_pass_twiddles(p->device, *fe, &c.pass_tw, dry))) return st;
        if ((st = get_two_level(p->device, n, &c.tw, dry))) return st;
        if ((st = get_two_level(p->device, L, &c.tw4, dry))) return st;
        l.in_buf = B_IN;
        l.out_buf = B_SCR;
        l.first = true;
        l.grid = c.n_tiles;
        p->launches.push_back(l);
        return tile((int)N2, PAT_ROW, PAT_COL, 0, B_SCR, 0, B_OUT, 0, N1, batch, 1, 0, L, N2, 1, 0, out_dist / 2, 1, N1,
                    TW_NONE, 0, false, true);
    }
};

int64_t dflt(int64_t v, int64_t d) { return v ? v : d; }

// Whether a launch may write the caller's output before every read of the
// caller's input has finished: some launch reads B_IN at or after the first
// launch that writes B_OUT (launches on one stream run in order).
bool output_before_input_done(const std::vector<Launch> &launches)
{
    size_t first_write = launches.size();
    for (size_t i = 0; i < launches.size(); i++)
        if (launches[i].writes(B_OUT)) {
            first_write = i;
            break;
        }
    for (size_t i = first_write; i < launches.size(); i++)
        if (launches[i].reads(B_IN)) return true;
    return false;
}

int plan_via_dense_copy(gfft_plan_s *p, bool dry, const PlanOpts *forced, const gfft_desc &dense, bool input_side,
                        const RepackArgs &rp);

int build_plan(gfft_plan_s *p, bool dry, const PlanOpts *forced = nullptr)
{
    const gfft_desc &d = p->d;
    Builder b{p, p->prec, csize_of(p->prec, p->store), dry, forced};
    b.store = p->store;
    if (p->store) b.opts = PlanOpts{};
    const size_t cs = b.csize, rs = cs / 2;
    const int64_t n = d.n, batch = d.batch;
    int st = GFFT_OK;
    switch (d.kind) {
    case GFFT_C2C: {
        const int64_t id = dflt(d.in_dist, n), od = dflt(d.out_dist, n);
        p->scale_div = id;
        p->amax_width = p->amax_pitch = p->amax_gain = n;
        p->in_bytes = ((batch - 1) * id + n) * cs;
        p->out_bytes = ((batch - 1) * od + n) * cs;
        p->out_has_gaps = od > n;
        p->in_step = id * (int64_t)cs;
        p->out_step = od * (int64_t)cs;
        st = b.c2c_1d(n, batch, B_IN, 0, id, B_OUT, 0, od, true, true);
        break;
    }
    case GFFT_R2C: {
        const bool half = d.flags & GFFT_FLAG_HALF;
        const int64_t nout = half ? n / 2 + 1 : n;
        const int64_t id = dflt(d.in_dist, n), od = dflt(d.out_dist, nout);
        p->in_bytes = ((batch - 1) * id + n) * rs;
        p->out_bytes = ((batch - 1) * od + nout) * cs;
        p->out_has_gaps = od > nout;
        p->in_step = id * (int64_t)rs;
        p->out_step = od * (int64_t)cs;
        p->dir_mask = 1;
        if (n == 1) {
            Launch l;
            l.type = LT_COPY;
            l.ca = {nullptr, nullptr, batch, id, od, 1, 0};
            p->launches.push_back(l);
            break;
        }
        if (id % 2) {
            gfft_desc dense = d;
            dense.in_dist = 0;
            return plan_via_dense_copy(p, dry, forced, dense, true, RepackArgs{nullptr, nullptr, n, batch, 1, id, 0, n, 0, (int32_t)rs, 0});
        }
        if (b.rows_fusable(n)) {
            st = b.real_rows(RR_R2C, n, batch, B_IN, 0, id / 2, B_OUT, 0, od, half);
            break;
        }
        // The fused path needs scratch for the whole batch.  The separate
        // unpack with an odd output distance needs the same; with an even
        // one it chunks the transform through at most GFFT_CHUNK_BYTES, so a
        // larger batch keeps that path.
        if (half && b.r2c_fusable(n) && b.opts.r2c_fuse && (od % 2 || batch * (n / 2) * (int64_t)cs <= b.chunk_bytes())) {
            st = b.r2c_fused(n, batch, id, od);
            break;
        }
        if (od % 2 == 0) {
            st = b.c2c_1d(n / 2, batch, B_IN, 0, id / 2, B_OUT, 0, od, true, true);
            if (!st) st = b.unpack(n, batch, B_OUT, od, B_OUT, od, half, false);
            break;
        }
        // An odd output distance (the dense n/2+1 half spectrum) would force
        // 8-byte stores in every tile pass: the n/2-point transform goes
        // densely into scratch instead, after the four-step ring, and the
        // unpack writes the output.
        {
            const int64_t L = n / 2;
            const size_t c0 = p->launches.size();
            st = b.c2c_1d(L, batch, B_IN, 0, id / 2, B_SCR, 0, L, true, true);
            if (st) break;
            const int64_t ring = b.scratch_elems;
            for (size_t i = c0; i < p->launches.size(); i++)
                if (p->launches[i].last && p->launches[i].out_buf == B_SCR) p->launches[i].out_off += ring * (int64_t)cs;
            st = b.unpack(n, batch, B_SCR, L, B_OUT, od, half, false);
            if (st) break;
            p->launches.back().in_off = ring * (int64_t)cs;
            b.scratch_elems = ring + L * batch;
        }
        break;
    }
    case GFFT_C2R: {
        const int64_t nin = n / 2 + 1;
        const int64_t id = dflt(d.in_dist, nin), od = dflt(d.out_dist, n);
        p->in_bytes = ((batch - 1) * id + nin) * cs;
        p->out_bytes = ((batch - 1) * od + n) * rs;
        p->out_has_gaps = od > n;
        p->in_step = id * (int64_t)cs;
        p->out_step = od * (int64_t)rs;
        p->dir_mask = 2;
        if (n == 1) {
            Launch l;
            l.type = LT_COPY;
            l.ca = {nullptr, nullptr, batch, id, od, 0, 1};
            p->launches.push_back(l);
            break;
        }
        if (od % 2) {
            gfft_desc dense = d;
            dense.out_dist = 0;
            return plan_via_dense_copy(p, dry, forced, dense, false, RepackArgs{nullptr, nullptr, n, batch, 1, n, 0, od, 0, (int32_t)rs, 0});
        }
        if (b.rows_fusable(n)) {
            st = b.real_rows(RR_C2R, n, batch, B_IN, 0, id, B_OUT, 0, od / 2, false);
            break;
        }
        if (b.c2r_fusable(n) && b.opts.r2c_fuse) {
            st = b.c2r_fused(n, batch, id, od);
            break;
        }
        const int64_t L = n / 2;
        b.need_scratch(L * batch);
        st = b.unpack(n, batch, B_IN, id, B_SCR, L, false, true);
        if (!st) st = b.c2c_1d(L, batch, B_SCR, 0, L, B_OUT, 0, od / 2, true, true);
        // scratch is shared with a four-step c2c: the pack target must not
        // overlap the four-step ring, so offset it past the ring.
        if (!st && L > max_len(p->prec, PAT_ROW, PAT_ROW)) {
            const int64_t ring = b.scratch_elems;
            b.scratch_elems = ring + L * batch;
            for (auto &l : p->launches) {
                if (l.type == LT_PACK) l.out_off = ring * (int64_t)cs;
                if (l.type == LT_TILE && l.first) l.in_off += ring * (int64_t)cs;
            }
        }
        break;
    }
    case GFFT_DIT: {
        const bool half = d.flags & GFFT_FLAG_HALF;
        const int64_t nin = std::max<int64_t>(1, n / 2), nout = half ? n / 2 + 1 : n;
        const int64_t id = dflt(d.in_dist, nin), od = dflt(d.out_dist, nout);
        p->in_bytes = ((batch - 1) * id + nin) * cs;
        p->out_bytes = ((batch - 1) * od + nout) * cs;
        p->out_has_gaps = od > nout;
        p->in_step = id * (int64_t)cs;
        p->out_step = od * (int64_t)cs;
        p->dir_mask = 1;
        st = b.unpack(n, batch, B_IN, id, B_OUT, od, half, false);
        break;
    }
    case GFFT_COLS: {
        const int64_t rows = n, cols = d.n2;
        const int64_t ip = dflt(d.in_stride, cols), op = dflt(d.out_stride, cols);
        if (ip < cols || op < cols) return fail(GFFT_ERR_INVALID_ARG, "row pitch smaller than the column count");
        const int64_t id = dflt(d.in_dist, rows * ip), od = dflt(d.out_dist, rows * op);
        p->scale_div = id;
        p->amax_rows = p->amax_gain = rows;
        p->amax_width = cols;
        p->amax_pitch = ip;
        p->in_bytes = ((batch - 1) * id + (rows - 1) * ip + cols) * cs;
        p->out_bytes = ((batch - 1) * od + (rows - 1) * op + cols) * cs;
        p->out_has_gaps = op > cols || od > rows * op;
        p->in_step = ip * (int64_t)cs;
        p->out_step = op * (int64_t)cs;
        p->in_step2 = id * (int64_t)cs;
        p->out_step2 = od * (int64_t)cs;
        st = b.columns(rows, cols, batch, B_IN, ip, id, B_OUT, op, od, true, true);
        break;
    }
    case GFFT_C2C_2D: {
        const int64_t W = n, H = d.n2;
        const int64_t ip = dflt(d.in_stride, W), op = dflt(d.out_stride, W);
        if (ip < W || op < W) return fail(GFFT_ERR_INVALID_ARG, "row pitch smaller than the width");
        const int64_t id = dflt(d.in_dist, H * ip), od = dflt(d.out_dist, H * op);
        p->scale_div = id;
        p->amax_rows = H;
        p->amax_width = W;
        p->amax_pitch = ip;
        p->amax_gain = W * H;
        p->in_bytes = ((batch - 1) * id + (H - 1) * ip + W) * cs;
        p->out_bytes = ((batch - 1) * od + (H - 1) * op + W) * cs;
        p->out_has_gaps = op > W || od > H * op;
        p->in_step = ip * (int64_t)cs;
        p->out_step = op * (int64_t)cs;
        p->in_step2 = id * (int64_t)cs;
        p->out_step2 = od * (int64_t)cs;
        if (W > max_len(p->prec, PAT_ROW, PAT_ROW, p->store))
            return fail(GFFT_ERR_UNSUPPORTED, "2D width %lld exceeds the single-pass row limit", (long long)W);
        if (W == 1) {
            st = b.columns(H, 1, batch, B_IN, ip, id, B_OUT, op, od, true, true);
            break;
        }
        // rows: H rows of batch matrices (two outer levels)
        st = b.tile((int)W, PAT_ROW, PAT_ROW, 0, B_IN, 0, B_OUT, 0, H, batch, 1, 0, id, ip, 1, 0, od, op, 1,
                    TW_NONE, 0, true, H == 1);
        if (!st && H > 1) st = b.columns(H, W, batch, B_OUT, op, od, B_OUT, op, od, false, true);
        break;
    }
    case GFFT_R2C_2D: {
        const int64_t M = n, N = d.n2;
        const int64_t ip = dflt(d.in_stride, M), op = dflt(d.out_stride, M);
        if (ip < M || op < M) return fail(GFFT_ERR_INVALID_ARG, "row pitch smaller than the width");
        const int64_t id = dflt(d.in_dist, N * ip), od = dflt(d.out_dist, N * op);
        p->in_bytes = ((batch - 1) * id + (N - 1) * ip + M) * rs;
        p->out_bytes = ((batch - 1) * od + (N - 1) * op + M) * cs;
        p->out_has_gaps = op > M || od > N * op;
        p->in_step = ip * (int64_t)rs;
        p->out_step = op * (int64_t)cs;
        p->in_step2 = id * (int64_t)rs;
        p->out_step2 = od * (int64_t)cs;
        p->dir_mask = 1;
        if (M == 1) {
            // a single real column: complex column FFT of (x, 0)
            return fail(GFFT_ERR_UNSUPPORTED, "R2C_2D needs width >= 2");
        }
        if ((ip % 2) || (id % 2)) {
            gfft_desc dense = d;
            dense.in_stride = dense.in_dist = 0;
            return plan_via_dense_copy(p, dry, forced, dense, true, RepackArgs{nullptr, nullptr, M, N, batch, ip, id, M, N * M, (int32_t)rs, 0});
        }
        const int64_t half = M / 2;
        // row r2c (half spectrum, M/2+1 bins) on every row of every matrix
        if (b.rows_fusable(M)) {
            for (int64_t m = 0; m < batch && !st; m++)
                st = b.real_rows(RR_R2C, M, N, B_IN, m * (id / 2), ip / 2, B_OUT, m * od, op, true);
        } else if (batch == 1) {
            st = b.c2c_1d(half, N, B_IN, 0, ip / 2, B_OUT, 0, op, true, true);
            if (!st) st = b.unpack(M, N, B_OUT, op, B_OUT, op, true, false);
        } else {
            for (int64_t m = 0; m < batch && !st; m++) {
                const size_t before = p->launches.size();
                st = b.c2c_1d(half, N, B_IN, m * (id / 2), ip / 2, B_OUT, m * od, op, true, true);
                if (st) break;
                st = b.unpack(M, N, B_OUT, op, B_OUT, op, true, false);
                if (st) break;
                p->launches.back().in_off = p->launches.back().out_off = m * od * (int64_t)cs;
                (void)before;
            }
        }
        const int64_t vcols = std::min<int64_t>(M / 2 + 1, M);
        if (!st && N > 1) st = b.columns(N, vcols, batch, B_OUT, op, od, B_OUT, op, od, true, true);
        if (!st && vcols < M) {
            Launch l;
            l.type = LT_HFILL;
            l.ha = {nullptr, M, N, vcols, op, od, batch};
            l.in_buf = l.out_buf = B_OUT;
            p->launches.push_back(l);
        }
        break;
    }
    case GFFT_C2R_2D: {
        // inverse of R2C_2D: the n/2+1 non-redundant columns of the spectrum
        // (height N, input pitch in complex) are inverse-transformed into
        // scratch region B; each row is then a half spectrum, turned into M
        // reals by the c2r (pack + inverse c2c) into the output (pitch in scalars)
        const int64_t M = n, N = d.n2, vcols = M / 2 + 1, L = M / 2;
        const int64_t ip = dflt(d.in_stride, M), op = dflt(d.out_stride, M);
        const int64_t id = dflt(d.in_dist, N * ip), od = dflt(d.out_dist, N * op);
        p->in_bytes = ((batch - 1) * id + (N - 1) * ip + vcols) * cs;
        p->out_bytes = ((batch - 1) * od + (N - 1) * op + M) * rs;
        p->out_has_gaps = op > M || od > N * op;
        p->in_step = ip * (int64_t)cs;
        p->out_step = op * (int64_t)rs;
        p->in_step2 = id * (int64_t)cs;
        p->out_step2 = od * (int64_t)rs;
        p->dir_mask = 2;
        if (M < 2) return fail(GFFT_ERR_UNSUPPORTED, "C2R_2D needs width >= 2");
        if ((op % 2) || (od % 2)) {
            gfft_desc dense = d;
            dense.out_stride = dense.out_dist = 0;
            return plan_via_dense_copy(p, dry, forced, dense, false, RepackArgs{nullptr, nullptr, M, N, batch, M, N * M, op, od, (int32_t)rs, 0});
        }
        const int64_t bsz = N * vcols;  // region B: batch x N x vcols, row pitch vcols
        const size_t c0 = p->launches.size();
        st = b.columns(N, vcols, batch, B_IN, ip, id, B_SCR, vcols, bsz, true, true);
        const size_t c1 = p->launches.size();
        std::vector<size_t> packs;
        const bool uniform = od == N * op;
        for (int64_t m = 0; m < (uniform ? 1 : batch) && !st; m++) {
            const int64_t rowsn = uniform ? N * batch : N;
            if (b.rows_fusable(M)) {
                // pack + inverse row FFT in one pass, region B -> output
                st = b.real_rows(RR_C2R, M, rowsn, B_SCR, m * bsz, vcols, B_OUT, m * (od / 2), op / 2, false);
                packs.push_back(p->launches.size() - 1);
                continue;
            }
            st = b.unpack(M, rowsn, B_SCR, vcols, B_OUT, op / 2, false, true);
            if (st) break;
            packs.push_back(p->launches.size() - 1);
            p->launches.back().in_off = m * bsz * (int64_t)cs;
            p->launches.back().out_off = m * (od / 2) * (int64_t)cs;
            st = b.c2c_1d(L, rowsn, B_OUT, m * (od / 2), op / 2, B_OUT, m * (od / 2), op / 2, true, true);
        }
        if (st) break;
        // region B goes after everything else the plan keeps in scratch (the
        // column sub-pass intermediate, the four-step ring), which are only
        // live before / after it is written and read
        const int64_t ring = b.scratch_elems;
        // the launches that write region B: the single column pass or the
        // second sub-pass (the last launch), or every copy of a height-1 plane
        for (size_t i = c0; i < c1; i++) {
            Launch &l = p->launches[i];
            if (i + 1 == c1 || l.type == LT_COPY) {
                if (l.out_buf != B_SCR) return fail(GFFT_ERR_UNSUPPORTED, "internal: C2R_2D column output");
                l.out_off += ring * (int64_t)cs;
            }
        }
        for (size_t i : packs) p->launches[i].in_off += ring * (int64_t)cs;
        b.scratch_elems = ring + bsz * batch;
        break;
    }
    default:
        return fail(GFFT_ERR_INVALID_ARG, "unknown transform kind %d", d.kind);
    }
    if (st) return st;
    b.scratch_elems = std::max<int64_t>(b.scratch_elems, 0);
    p->scratch_bytes = (size_t)b.scratch_elems * cs;
    p->overlap_hazard = output_before_input_done(p->launches);
    return GFFT_OK;
}

// A real side whose pitch or distance is odd.  The kernels move reals as
// complex pairs, so every real row must start at an even scalar offset; the
// transform is planned instead on a dense copy of that side held in scratch
// (`dense`: the descriptor with that side's pitches at their defaults).  For
// an input side a repack launch fills the copy first; for an output side the
// last launch copies the result out of it.  genFFT's RealFFT2D::forward takes
// any scalar in_stride (include/genFFT/FFTReal.h:83, rows read at :144-165).
int plan_via_dense_copy(gfft_plan_s *p, bool dry, const PlanOpts *forced, const gfft_desc &dense, bool input_side,
                        const RepackArgs &rp)
{
    gfft_plan_s inner;
    inner.d = dense;
    inner.prec = p->prec;
    inner.store = p->store;
    inner.device = p->device;
    inner.num_cus = p->num_cus;
    int st = build_plan(&inner, dry, forced);
    if (st) return st;
    // the copy sits after everything the inner plan keeps in scratch
    const int64_t region = (int64_t)((inner.scratch_bytes + 255) & ~(size_t)255);
    const int moved = input_side ? B_IN : B_OUT;
    Launch r;
    r.type = LT_REPACK;
    r.rp = rp;
    r.in_buf = input_side ? B_IN : B_SCR;
    r.in_off = input_side ? 0 : region;
    r.out_buf = input_side ? B_SCR : B_OUT;
    r.out_off = input_side ? region : 0;
    p->launches.clear();
    if (input_side) p->launches.push_back(r);
    for (Launch l : inner.launches) {
        if (l.in_buf == moved) {
            l.in_buf = B_SCR;
            l.in_off += region;
        }
        if (l.out_buf == moved) {
            l.out_buf = B_SCR;
            l.out_off += region;
        }
        if (l.type == LT_DUAL && l.in_buf_b == moved) {
            l.in_buf_b = B_SCR;
            l.in_off_b += region;
        }
        if (l.type == LT_DUAL && l.out_buf_b == moved) {
            l.out_buf_b = B_SCR;
            l.out_off_b += region;
        }
        p->launches.push_back(l);
    }
    if (!input_side) p->launches.push_back(r);
    p->scratch_bytes = (size_t)region + (input_side ? inner.in_bytes : inner.out_bytes);
    // The caller's input is read only before (input side) or the output
    // written only after (output side) every other launch: overlapping
    // buffers cannot be corrupted.
    p->overlap_hazard = false;
    return GFFT_OK;
}

// Plan-time bounds proof: the largest byte offset each launch can touch,
// from the same index formulas the kernels use, must lie inside its buffer.
// A planner mistake becomes an error here instead of a device fault.
int validate(const gfft_plan_s *p)
{
    const int64_t cs = (int64_t)csize_of(p->prec, p->store), rs = cs / 2;
    const int64_t ext[3] = {(int64_t)p->in_bytes, (int64_t)p->out_bytes, (int64_t)p->scratch_bytes};
    auto inside = [&](int buf, int64_t off, int64_t last_byte, const char *what) {
        if (off < 0 || last_byte < 0 || off + last_byte > ext[buf])
            return fail(GFFT_ERR_UNSUPPORTED,
                        "internal: %s reaches byte %lld of buffer %d (extent %lld)", what,
                        (long long)(off + last_byte), buf, (long long)ext[buf]);
        return GFFT_OK;
    };
    int idx = 0;
    // one tile pass: every sequence it loads and stores lies inside its buffers
    auto tile_inside = [&](const KernelEntry *k, const TileArgs &a, int in_buf, int64_t in_off, int out_buf,
                           int64_t out_off) {
        const int L = k->L;
        const int64_t n_outer2 = a.n_tiles / ((int64_t)a.tiles_inner * a.n_outer1);
        if ((int64_t)a.tiles_inner * a.n_outer1 * n_outer2 != a.n_tiles)
            return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d tile count", idx);
        const int64_t li = (n_outer2 - 1) * a.in_outer2 + (int64_t)(a.n_outer1 - 1) * a.in_outer +
                           (int64_t)(a.n_inner - 1) * a.in_inner + (int64_t)(L - 1) * a.in_elem;
        const int64_t lo = (n_outer2 - 1) * a.out_outer2 + (int64_t)(a.n_outer1 - 1) * a.out_outer +
                           (int64_t)(a.n_inner - 1) * a.out_inner + (int64_t)(L - 1) * a.out_elem;
        int st;
        if ((st = inside(in_buf, in_off, (li + 1) * cs, "tile load"))) return st;
        if ((st = inside(out_buf, out_off, (lo + 1) * cs, "tile store"))) return st;
        if (k->lds_bytes > 160 * 1024) return fail(GFFT_ERR_UNSUPPORTED, "internal: LDS");
        return GFFT_OK;
    };
    for (const Launch &l : p->launches) {
        int st = GFFT_OK;
        idx++;
        switch (l.type) {
        case LT_TILE:
            if ((st = tile_inside(l.k, l.ta, l.in_buf, l.in_off, l.out_buf, l.out_off))) return st;
            break;
        case LT_DUAL: {
            if ((st = tile_inside(l.k, l.ta, l.in_buf, l.in_off, l.out_buf, l.out_off))) return st;
            if ((st = tile_inside(l.kb, l.tb, l.in_buf_b, l.in_off_b, l.out_buf_b, l.out_off_b))) return st;
            if (!l.dual || l.dual->lds_bytes > 160 * 1024 || l.split < 1 || l.grid <= l.split ||
                l.split > l.ta.n_tiles || l.grid - l.split > l.tb.n_tiles)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d dual split", idx);
            // the two roles never touch the same bytes of one buffer
            auto span = [&](const KernelEntry *k, const TileArgs &a, bool out, int64_t off) {
                const int64_t n_outer2 = a.n_tiles / ((int64_t)a.tiles_inner * a.n_outer1);
                const int64_t last = out ? (n_outer2 - 1) * a.out_outer2 + (int64_t)(a.n_outer1 - 1) * a.out_outer +
                                               (int64_t)(a.n_inner - 1) * a.out_inner + (int64_t)(k->L - 1) * a.out_elem
                                         : (n_outer2 - 1) * a.in_outer2 + (int64_t)(a.n_outer1 - 1) * a.in_outer +
                                               (int64_t)(a.n_inner - 1) * a.in_inner + (int64_t)(k->L - 1) * a.in_elem;
                return std::make_pair(off, off + (last + 1) * cs);
            };
            const auto aw = span(l.k, l.ta, true, l.out_off), br = span(l.kb, l.tb, false, l.in_off_b);
            if (l.out_buf == l.in_buf_b && aw.first < br.second && br.first < aw.second)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d roles overlap", idx);
            break;
        }
        case LT_UNPACK: {
            const UnpackArgs &a = l.ua;
            const int64_t nin = std::max<int64_t>(1, a.n / 2), nout = a.half ? a.n / 2 + 1 : std::max<int64_t>(1, a.n);
            if ((st = inside(l.in_buf, l.in_off, ((a.batch - 1) * a.in_dist + nin) * cs, "unpack load"))) return st;
            if ((st = inside(l.out_buf, l.out_off, ((a.batch - 1) * a.out_dist + nout) * cs, "unpack store")))
                return st;
            break;
        }
        case LT_PACK: {
            const UnpackArgs &a = l.ua;
            if ((st = inside(l.in_buf, l.in_off, ((a.batch - 1) * a.in_dist + a.n / 2 + 1) * cs, "pack load")))
                return st;
            if ((st = inside(l.out_buf, l.out_off, ((a.batch - 1) * a.out_dist + a.n / 2) * cs, "pack store")))
                return st;
            break;
        }
        case LT_RROWS: {
            const TileArgs &a = l.rr.t;
            const int64_t L = l.k->L;
            const int64_t nin = l.rr_mode == RR_R2C ? L : L + 1;
            const int64_t nout = l.rr_mode == RR_C2R ? L : (l.rr.half ? L + 1 : 2 * L);
            if ((st = inside(l.in_buf, l.in_off, ((int64_t)(a.n_inner - 1) * a.in_inner + nin) * cs, "real rows load")))
                return st;
            if ((st = inside(l.out_buf, l.out_off, ((int64_t)(a.n_inner - 1) * a.out_inner + nout) * cs,
                             "real rows store")))
                return st;
            if (l.k->lds_bytes > 160 * 1024) return fail(GFFT_ERR_UNSUPPORTED, "internal: LDS");
            break;
        }
        case LT_C2R: {
            const C2rArgs &a = l.cr;
            const int64_t nb = a.n_tiles / a.tiles_per, L = (int64_t)l.k->L * a.n2;
            if ((int64_t)a.tiles_per * nb != a.n_tiles || (int64_t)a.tiles_per * l.k->M != a.n2)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d c2r tiles", idx);
            if ((st = inside(l.in_buf, l.in_off, ((nb - 1) * a.in_dist + L + 1) * cs, "c2r bin load"))) return st;
            if ((st = inside(l.out_buf, l.out_off, nb * L * cs, "c2r column store"))) return st;
            if (l.k->lds_bytes > 160 * 1024) return fail(GFFT_ERR_UNSUPPORTED, "internal: LDS");
            break;
        }
        case LT_R2C: {
            const R2cArgs &a = l.ra;
            const int64_t nb = a.n_tiles / a.tiles_per, L = (int64_t)a.n1 * l.k->L;
            if ((int64_t)a.tiles_per * nb != a.n_tiles || (int64_t)a.tiles_per * l.k->M != a.n1)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d r2c tiles", idx);
            if ((st = inside(l.in_buf, l.in_off, ((nb - 1) * a.in_dist + L) * cs, "r2c row load"))) return st;
            if ((st = inside(l.out_buf, l.out_off, ((nb - 1) * a.out_dist + L + 1) * cs, "r2c bin store"))) return st;
            if (l.k->lds_bytes > 160 * 1024) return fail(GFFT_ERR_UNSUPPORTED, "internal: LDS");
            break;
        }
        case LT_HFILL: {
            const HFillArgs &a = l.ha;
            if ((st = inside(l.out_buf, l.out_off,
                             ((a.batch - 1) * a.dist + (a.N - 1) * a.stride + a.M) * cs, "hfill")))
                return st;
            break;
        }
        case LT_REPACK: {
            const RepackArgs &a = l.rp;
            if (a.width < 1 || a.rows < 1 || a.batch < 1)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d repack shape", idx);
            if (a.elem_bytes != 4 && a.elem_bytes != 8)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d repack element size", idx);
            if ((st = inside(l.in_buf, l.in_off,
                             ((a.batch - 1) * a.in_dist + (a.rows - 1) * a.in_pitch + a.width) * a.elem_bytes,
                             "repack load")))
                return st;
            if ((st = inside(l.out_buf, l.out_off,
                             ((a.batch - 1) * a.out_dist + (a.rows - 1) * a.out_pitch + a.width) * a.elem_bytes,
                             "repack store")))
                return st;
            break;
        }
        case LT_COPY: {
            const CopyArgs &a = l.ca;
            const int64_t isz = a.real_in ? rs : cs, osz = a.real_out ? rs : cs;
            if ((st = inside(l.in_buf, l.in_off, ((a.batch - 1) * a.in_dist + 1) * isz, "copy load"))) return st;
            if ((st = inside(l.out_buf, l.out_off, ((a.batch - 1) * a.out_dist + 1) * osz, "copy store")))
                return st;
            break;
        }
        case LT_MIX:
        case LT_MIX_RESET:
            return fail(GFFT_ERR_UNSUPPORTED, "internal: launch %d belongs to batch plans", idx);
        }
    }
    return GFFT_OK;
}

int describe_launches(const std::vector<Launch> &launches, const std::string &head, char *buf, size_t len);

int describe(const gfft_plan_s *p, char *buf, size_t len)
{
    char line[256];
    snprintf(line, sizeof(line), "plan in_bytes=%zu out_bytes=%zu scratch_bytes=%zu\n", p->in_bytes,
             p->out_bytes, p->scratch_bytes);
    return describe_launches(p->launches, line, buf, len);
}

int describe_launches(const std::vector<Launch> &launches, const std::string &head, char *buf, size_t len)
{
    std::string s = head;
    char line[1024];
    const char *bn[3] = {"in", "out", "scratch"};
    for (const Launch &l : launches) {
        switch (l.type) {
        case LT_TILE: {
            const TileArgs &a = l.ta;
            snprintf(line, sizeof(line),
                     "tile L=%d M=%d NT=%d layout=%d inp=%d outp=%d lds=%d pipe=%d grid=%lld rad=%d,%d,%d,%d tiles=%d "
                     "n_inner=%d tiles_inner=%d n_outer1=%d in=%s+%lld out=%s+%lld "
                     "in_strides=%lld,%lld,%lld,%lld out_strides=%lld,%lld,%lld,%lld tw=%d first=%d last=%d vin=%d vout=%d runs=%d\n",
                     l.k->L, l.k->M, l.k->NT, l.k->layout, l.k->inp, l.k->outp, l.k->lds_bytes, l.k->persistent,
                     (long long)l.grid, l.k->rad[0],
                     l.k->nrad > 1 ? l.k->rad[1] : 0, l.k->nrad > 2 ? l.k->rad[2] : 0,
                     l.k->nrad > 3 ? l.k->rad[3] : 0, a.n_tiles, a.n_inner, a.tiles_inner, a.n_outer1,
                     bn[l.in_buf], (long long)l.in_off, bn[l.out_buf], (long long)l.out_off,
                     (long long)a.in_outer2, (long long)a.in_outer, (long long)a.in_inner, (long long)a.in_elem,
                     (long long)a.out_outer2, (long long)a.out_outer, (long long)a.out_inner,
                     (long long)a.out_elem, a.tw_mode, (int)l.first, (int)l.last, (int)l.vin_ok, (int)l.vout_ok, a.runs);
            break;
        }
        case LT_DUAL: {
            auto role = [&](const KernelEntry *k, const TileArgs &a, int ib, int64_t io, int ob, int64_t oo) {
                char r[384];
                snprintf(r, sizeof(r), "L=%d M=%d inp=%d outp=%d rad=%d,%d,%d,%d tiles=%d n_inner=%d tiles_inner=%d "
                         "n_outer1=%d in=%s+%lld out=%s+%lld tw=%d runs=%d", k->L, k->M, k->inp, k->outp, k->rad[0],
                         k->nrad > 1 ? k->rad[1] : 0, k->nrad > 2 ? k->rad[2] : 0, k->nrad > 3 ? k->rad[3] : 0,
                         a.n_tiles, a.n_inner, a.tiles_inner, a.n_outer1, bn[ib], (long long)io, bn[ob], (long long)oo,
                         a.tw_mode, a.runs);
                return std::string(r);
            };
            snprintf(line, sizeof(line), "dual NT=%d lds=%d grid=%lld split=%lld | %s | %s\n", l.dual->NT,
                     l.dual->lds_bytes, (long long)l.grid, (long long)l.split,
                     role(l.k, l.ta, l.in_buf, l.in_off, l.out_buf, l.out_off).c_str(),
                     role(l.kb, l.tb, l.in_buf_b, l.in_off_b, l.out_buf_b, l.out_off_b).c_str());
            break;
        }
        case LT_UNPACK:
        case LT_PACK:
            snprintf(line, sizeof(line), "%s n=%lld batch=%lld in=%s+%lld out=%s+%lld in_dist=%lld out_dist=%lld "
                     "half=%d\n", l.type == LT_PACK ? "pack" : "unpack", (long long)l.ua.n, (long long)l.ua.batch,
                     bn[l.in_buf], (long long)l.in_off, bn[l.out_buf], (long long)l.out_off,
                     (long long)l.ua.in_dist, (long long)l.ua.out_dist, l.ua.half);
            break;
        case LT_RROWS:
            snprintf(line, sizeof(line), "real_rows mode=%s L=%d M=%d NT=%d rows=%d tiles=%d half=%d in=%s+%lld "
                     "out=%s+%lld in_row=%lld out_row=%lld\n", l.rr_mode == RR_R2C ? "r2c" : "c2r", l.k->L, l.k->M,
                     l.k->NT, l.rr.t.n_inner, l.rr.t.n_tiles, l.rr.half, bn[l.in_buf], (long long)l.in_off,
                     bn[l.out_buf], (long long)l.out_off, (long long)l.rr.t.in_inner, (long long)l.rr.t.out_inner);
            break;
        case LT_C2R:
            snprintf(line, sizeof(line), "c2r_pack_cols L=%d M=%d NT=%d n2=%d tiles=%d in=%s+%lld out=%s+%lld "
                     "in_dist=%lld\n", l.k->L, l.k->M, l.k->NT, l.cr.n2, l.cr.n_tiles, bn[l.in_buf], (long long)l.in_off,
                     bn[l.out_buf], (long long)l.out_off, (long long)l.cr.in_dist);
            break;
        case LT_R2C:
            snprintf(line, sizeof(line), "r2c_rows_unpack R=%d M=%d NT=%d n1=%d tiles=%d grid=%lld in=%s+%lld out=%s+%lld "
                     "in_dist=%lld out_dist=%lld\n", l.k->L, l.k->M, l.k->NT, l.ra.n1, l.ra.n_tiles, (long long)l.grid,
                     bn[l.in_buf],
                     (long long)l.in_off, bn[l.out_buf], (long long)l.out_off, (long long)l.ra.in_dist,
                     (long long)l.ra.out_dist);
            break;
        case LT_HFILL:
            snprintf(line, sizeof(line), "hfill M=%lld N=%lld vcols=%lld stride=%lld dist=%lld batch=%lld\n",
                     (long long)l.ha.M, (long long)l.ha.N, (long long)l.ha.vcols, (long long)l.ha.stride,
                     (long long)l.ha.dist, (long long)l.ha.batch);
            break;
        case LT_COPY:
            snprintf(line, sizeof(line), "copy batch=%lld in=%s+%lld out=%s+%lld\n", (long long)l.ca.batch,
                     bn[l.in_buf], (long long)l.in_off, bn[l.out_buf], (long long)l.out_off);
            break;
        case LT_REPACK:
            snprintf(line, sizeof(line), "repack width=%lld rows=%lld batch=%lld in=%s+%lld out=%s+%lld "
                     "in_pitch=%lld in_dist=%lld out_pitch=%lld out_dist=%lld\n", (long long)l.rp.width,
                     (long long)l.rp.rows, (long long)l.rp.batch, bn[l.in_buf], (long long)l.in_off, bn[l.out_buf],
                     (long long)l.out_off, (long long)l.rp.in_pitch, (long long)l.rp.in_dist,
                     (long long)l.rp.out_pitch, (long long)l.rp.out_dist);
            break;
        case LT_MIX:
            snprintf(line, sizeof(line), "mix tasks=%d grid=%lld\n", l.mix.n_tasks, (long long)l.grid);
            break;
        case LT_MIX_RESET:
            snprintf(line, sizeof(line), "mix_reset\n");
            break;
        }
        s += line;
    }
    if (s.size() + 1 > len) return fail(GFFT_ERR_INVALID_ARG, "describe buffer too small (%zu bytes needed)", s.size() + 1);
    memcpy(buf, s.c_str(), s.size() + 1);
    return GFFT_OK;
}

// Executes the plan's launches in order on stream s.  With `events`, records
// events[2i] / events[2i+1] on s around launch i (per-kernel timing).
// With `scale` (16-bit storage plans), the tile launches that read the plan's
// input multiply transform b's values by scale[b] on their store.
int run_launches(const std::vector<Launch> &launches, int prec, int xcd_remap, const void *in, void *out,
                 void *scratch, int direction, hipStream_t s, hipEvent_t *events = nullptr,
                 const float *scale = nullptr, int64_t scale_div = 1, hipEvent_t stop = nullptr,
                 bool *stopped = nullptr);

// Whether this call's buffers overlap in a way the plan's launch order can
// corrupt (see gfft_plan_s::overlap_hazard).
bool needs_stage(const gfft_plan_s *p, const void *in, const void *out)
{
    if (!p->overlap_hazard) return false;
    const uintptr_t i0 = (uintptr_t)in, o0 = (uintptr_t)out;
    if (!(i0 < o0 + p->out_bytes && o0 < i0 + p->in_bytes)) return false;
    return !(i0 == o0 && p->in_step == p->out_step && p->in_step2 == p->out_step2);
}

int run(gfft_plan_s *p, const void *in, void *out, int direction, hipStream_t s,
        hipEvent_t *events = nullptr, const float *scale = nullptr)
{
    if (p->scratch_bytes && !p->scratch)
        return fail(GFFT_ERR_INVALID_ARG, "plan created with GFFT_FLAG_NO_WORKSPACE holds no workspace: "
                                          "run it with gfft_execute_noalloc");
    const bool stage = needs_stage(p, in, out);
    if (!p->scratch && !stage)  // no shared workspace: calls on any streams are independent
        return run_launches(p->launches, p->prec, p->xcd_remap, in, out, p->scratch, direction, s, events, scale,
                            p->scale_div);
    std::lock_guard<std::mutex> lk(p->order.mu);
    if (stage && !p->stage) {
        hipError_t e = hipMalloc(&p->stage, p->in_bytes);
        if (e != hipSuccess) {
            p->stage = nullptr;
            return fail(GFFT_ERR_ALLOC, "staging buffer of %zu bytes for overlapping in/out: %s", p->in_bytes,
                        hipGetErrorString(e));
        }
    }
    HIP_TRY(p->order.before(s));
    if (stage) {
        // the whole input first, so no launch reads what another has overwritten
        HIP_TRY(hipMemcpyAsync(p->stage, in, p->in_bytes, hipMemcpyDeviceToDevice, s));
        in = p->stage;
    }
    bool stopped = false;
    int st = run_launches(p->launches, p->prec, p->xcd_remap, in, out, p->scratch, direction, s, events, scale,
                          p->scale_div, p->order.stop_event(s), &stopped);
    if (st) return st;
    if (stopped) p->order.mark(s);
    else HIP_TRY(p->order.after(s));
    return GFFT_OK;
}

int run_launches(const std::vector<Launch> &launches, int prec, int xcd_remap, const void *in, void *out,
                 void *scratch, int direction, hipStream_t s, hipEvent_t *events, const float *scale,
                 int64_t scale_div, hipEvent_t stop, bool *stopped)
{
    struct {
        int prec, xcd_remap;
    } pp{prec, xcd_remap}, *p = &pp;
    const bool inv = direction == GFFT_INVERSE;
    char *bufs[3] = {(char *)in, (char *)out, (char *)scratch};
    int li = 0;
    for (const Launch &L0 : launches) {
        Launch l = L0;
        // `stop` completes with the last launch when that is a tile, real-row,
        // fused r2c, unpack or Hermitian-fill kernel (*stopped says so); other
        // kinds leave it to the caller
        hipEvent_t tail = stop && &L0 == &launches.back() ? stop : nullptr;
        if (events) HIP_TRY(hipEventRecord(events[2 * li], s));
        char *ip = bufs[l.in_buf] + l.in_off;
        char *op = bufs[l.out_buf] + l.out_off;
        switch (l.type) {
        case LT_TILE: {
            TileArgs a = l.ta;
            a.in = ip;
            a.out = op;
            a.swap_in = inv && l.first;
            a.swap_out = inv && l.last;
            a.vec_in = l.vin_ok && (((uintptr_t)ip & 15) == 0);
            a.vec_out = l.vout_ok && (((uintptr_t)op & 15) == 0);
            // workgroups that each take one contiguous run of tiles keep the
            // hardware's block order (measured: c64 2^20 column pass 0.441 ->
            // 0.435 ms); interleaved tiles keep neighbours on one XCD
            a.xcd_remap = p->xcd_remap && a.runs >= 0;
            if (scale && l.in_buf == B_IN) {
                a.scale = scale;
                a.scale_div = scale_div;
                a.scale_off = l.in_off / (int64_t)csize_of(l.k->prec, l.k->store);
            }
            void *args[] = {&a};
            HIP_TRY(hipExtLaunchKernel((const void *)l.k->fn, dim3((unsigned)l.grid), dim3(l.k->NT), args, 0, s,
                                       nullptr, tail, 0));
            if (tail && stopped) *stopped = true;
            break;
        }
        case LT_DUAL: {
            auto role = [&](const TileArgs &t, int ib, int64_t io, int ob, int64_t oo, bool vi, bool vo, bool first,
                            bool last) {
                TileArgs a = t;
                a.in = bufs[ib] + io;
                a.out = bufs[ob] + oo;
                a.swap_in = inv && first;
                a.swap_out = inv && last;
                a.vec_in = vi && (((uintptr_t)a.in & 15) == 0);
                a.vec_out = vo && (((uintptr_t)a.out & 15) == 0);
                a.xcd_remap = p->xcd_remap && a.runs >= 0;
                return a;
            };
            TileArgs a = role(l.ta, l.in_buf, l.in_off, l.out_buf, l.out_off, l.vin_ok, l.vout_ok, l.first, l.last);
            TileArgs b = role(l.tb, l.in_buf_b, l.in_off_b, l.out_buf_b, l.out_off_b, l.vin_ok_b, l.vout_ok_b,
                              l.first_b, l.last_b);
            int split = (int)l.split;
            void *args[] = {&a, &b, &split};
            HIP_TRY(hipExtLaunchKernel((const void *)l.dual->fn, dim3((unsigned)l.grid), dim3(l.dual->NT), args, 0,
                                       s, nullptr, tail, 0));
            if (tail && stopped) *stopped = true;
            break;
        }
        case LT_UNPACK:
        case LT_PACK: {
            UnpackArgs a = l.ua;
            a.in = ip;
            a.out = op;
            HIP_TRY(l.type == LT_UNPACK ? launch_unpack(p->prec, a, s, tail) : launch_pack(p->prec, a, s));
            if (tail && stopped && l.type == LT_UNPACK) *stopped = true;
            break;
        }
        case LT_HFILL: {
            HFillArgs a = l.ha;
            a.out = op;
            HIP_TRY(launch_hfill(p->prec, a, s, tail));
            if (tail && stopped) *stopped = true;
            break;
        }
        case LT_RROWS: {
            RealRowsArgs a = l.rr;
            a.t.in = ip;
            a.t.out = op;
            a.t.vec_in = l.vin_ok && (((uintptr_t)ip & 15) == 0);
            a.t.vec_out = l.vout_ok && (((uintptr_t)op & 15) == 0);
            a.t.xcd_remap = p->xcd_remap;
            HIP_TRY(launch_real_rows(p->prec, l.k->L, l.rr_mode, a, l.grid, s, tail));
            if (tail && stopped) *stopped = true;
            break;
        }
        case LT_C2R: {
            C2rArgs a = l.cr;
            a.in = ip;
            a.out = op;
            a.xcd_remap = p->xcd_remap;
            HIP_TRY(launch_c2r_fused(p->prec, l.k->L, a, s));
            break;
        }
        case LT_R2C: {
            R2cArgs a = l.ra;
            a.in = ip;
            a.out = op;
            a.vec_in = l.vin_ok && (((uintptr_t)ip & 15) == 0);
            a.xcd_remap = p->xcd_remap;
            HIP_TRY(launch_r2c_fused(p->prec, a, s, tail));
            if (tail && stopped) *stopped = true;
            break;
        }
        case LT_COPY: {
            CopyArgs a = l.ca;
            a.in = ip;
            a.out = op;
            HIP_TRY(launch_copy(p->prec, a, s));
            break;
        }
        case LT_REPACK: {
            RepackArgs a = l.rp;
            a.in = ip;
            a.out = op;
            HIP_TRY(launch_repack(a, s));
            break;
        }
        case LT_MIX_RESET:
            HIP_TRY(launch_mix_reset(l.mix.counter, s));
            break;
        case LT_MIX: {
            MixArgs a = l.mix;
            a.in = ip;
            a.out = op;
            a.vec = l.vin_ok && (((uintptr_t)ip & 15) == 0) && (((uintptr_t)op & 15) == 0);
            HIP_TRY(launch_mix(p->prec, a, (int)l.grid, s));
            break;
        }
        }
        if (events) HIP_TRY(hipEventRecord(events[2 * li + 1], s));
        li++;
    }
    return GFFT_OK;
}

bool device_ok(int *dev)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return false;
    return hipGetDevice(dev) == hipSuccess;
}

// A plan's tables and workspace live on the device that was current when it
// was created; its launches must run there.
int on_plan_device(int plan_device)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return fail(GFFT_ERR_HIP, "hipGetDevice failed");
    if (cur != plan_device)
        return fail(GFFT_ERR_INVALID_ARG, "plan belongs to device %d but device %d is current", plan_device, cur);
    return GFFT_OK;
}

// Device address of noalloc_scratch on device `dev`, which must be current:
// resolved once per device (each GPU holds its own copy of the array at its
// own address) and kept.  Kernels and copies take this address, never the
// array's host-side name.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch_ptr(int dev, void **ptr)
{
    if (dev < 0 || dev >= kMaxDevices)
        return fail(GFFT_ERR_UNSUPPORTED, "device %d beyond the %d the static workspace table covers", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !a)
            return fail(GFFT_ERR_HIP, "hipGetSymbolAddress(noalloc_scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_noalloc_ptr[dev] = a;
    }
    *ptr = g_noalloc_ptr[dev];
    return GFFT_OK;
}

// A plan created without a workspace must fit the static one.
int check_noalloc_fit(const gfft_plan_s *p)
{
    if ((p->d.flags & GFFT_FLAG_NO_WORKSPACE) && p->scratch_bytes > (size_t)GFFT_NOALLOC_SCRATCH_BYTES)
        return fail(GFFT_ERR_UNSUPPORTED,
                    "this plan needs %zu bytes of workspace, more than the %lld bytes of the static workspace "
                    "(GFFT_FLAG_NO_WORKSPACE / gfft_execute_noalloc)",
                    p->scratch_bytes, (long long)GFFT_NOALLOC_SCRATCH_BYTES);
    return GFFT_OK;
}


// ---------------------------------------------------------------------------
// Uneven batches: transforms of mixed power-of-two lengths at arbitrary
// offsets, each with its own direction.
//   lengths 1 .. MixCfg::TOT   one persistent mixed-size launch (device work
//                              queue), preceded by a one-thread counter reset;
//   longer lengths            per length class, the single-pass or four-step
//                              tile launches with job tables along the
//                              transform index.
// The whole list is captured once into a HIP graph and replayed.  (Captured
// with the classes forked onto four side streams, the replay ran slower than
// one stream: the persistent class kernels each size their grid to the whole
// GPU and contend when they overlap.)
// ---------------------------------------------------------------------------
}  // namespace
}  // namespace gfft

struct gfft_batch_plan_s {
    int prec = 0, device = 0, num_cus = 256, xcd_remap = 1;
    int64_t count = 0;
    std::vector<int64_t> sizes, in_off, out_off;
    std::vector<int32_t> dirs;
    int64_t *d_job_in = nullptr, *d_job_out = nullptr;
    uint8_t *d_job_inv = nullptr;
    gfft::MixTask *d_tasks = nullptr;
    uint32_t *d_counter = nullptr;
    std::vector<gfft::Launch> launches;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    bool offsets_even = true;
    std::mutex mu;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *g_in = nullptr;
    void *g_out = nullptr;
    hipStream_t cap_stream = nullptr;
    std::map<int64_t, gfft_plan> per_job;  // one single-transform plan per length (reference loop)
    StreamOrder order;  // scratch, counter and the captured graph are shared by all calls
};

namespace gfft {
namespace {

int build_batch(gfft_batch_plan_s *p)
{
    const int prec = p->prec;
    const size_t cs = prec ? 16 : 8;
    const int64_t n = p->count;
    // sort jobs by length (stable: original order inside a class)
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return p->sizes[a] < p->sizes[b]; });
    std::vector<int64_t> jin(n), jout(n);
    std::vector<uint8_t> jinv(n);
    for (int64_t k = 0; k < n; k++) {
        const int64_t j = order[k];
        jin[k] = p->in_off[j];
        jout[k] = p->out_off[j];
        jinv[k] = p->dirs[j] == GFFT_INVERSE ? 1 : 0;
        p->offsets_even = p->offsets_even && (jin[k] % 2 == 0) && (jout[k] % 2 == 0);
    }
    int st;
    if ((st = upload((void **)&p->d_job_in, jin.data(), n * sizeof(int64_t)))) return st;
    if ((st = upload((void **)&p->d_job_out, jout.data(), n * sizeof(int64_t)))) return st;
    if ((st = upload((void **)&p->d_job_inv, jinv.data(), n))) return st;
    HIP_TRY(hipMalloc((void **)&p->d_counter, 256));
    HIP_TRY(hipMemset(p->d_counter, 0, 256));

    // classes
    const int mix_max = prec ? MixCfg<double>::MAXLG : MixCfg<float>::MAXLG;
    const int mix_tot = prec ? MixCfg<double>::TOT : MixCfg<float>::TOT;
    std::vector<MixTask> tasks;
    struct Cls {
        int lg;
        int64_t first, count;
    };
    std::vector<Cls> big;
    for (int64_t k = 0; k < n;) {
        int64_t e = k;
        while (e < n && p->sizes[order[e]] == p->sizes[order[k]]) e++;
        const int lg = ilog2(p->sizes[order[k]]);
        if (lg <= mix_max) {
            const int m = lg == 0 ? 64 : mix_m(mix_tot, 1 << lg);
            for (int64_t f = k; f < e; f += m) tasks.push_back(MixTask{lg, (int32_t)f, (int32_t)std::min<int64_t>(m, e - f), 0});
        } else {
            big.push_back(Cls{lg, k, e - k});
        }
        k = e;
    }
    // largest tiles first: better balance for the dynamic queue
    std::stable_sort(tasks.begin(), tasks.end(), [](const MixTask &a, const MixTask &b) { return a.lg > b.lg; });

    if (!tasks.empty()) {
        if ((st = upload((void **)&p->d_tasks, tasks.data(), tasks.size() * sizeof(MixTask)))) return st;
        Launch r;
        r.type = LT_MIX_RESET;
        r.mix.counter = p->d_counter;
        p->launches.push_back(r);
        Launch l;
        l.type = LT_MIX;
        l.mix.job_in = p->d_job_in;
        l.mix.job_out = p->d_job_out;
        l.mix.job_inv = p->d_job_inv;
        l.mix.tasks = p->d_tasks;
        l.mix.counter = p->d_counter;
        l.mix.n_tasks = (int32_t)tasks.size();
        for (int lg = 1; lg <= mix_max; lg++) {
            KernelEntry k{};
            k.prec = prec;
            k.L = 1 << lg;
            k.nrad = mix_radices(prec, lg, k.rad);
            if ((st = get_pass_twiddles(p->device, k, &l.mix.pass_tw[lg], false))) return st;
        }
        l.vin_ok = p->offsets_even && prec == 0;
        l.grid = std::min<int64_t>((int64_t)tasks.size(), mix_grid(prec, p->num_cus));
        p->launches.push_back(l);
    }

    // length classes above the mixed kernel: job tables along the transform
    // index, one class after another, the four-step classes sharing one
    // scratch region
    gfft_plan_s tmp;
    tmp.prec = prec;
    tmp.device = p->device;
    tmp.num_cus = p->num_cus;
    Builder b{&tmp, prec, cs, false};
    const int maxrow = max_len(prec, PAT_ROW, PAT_ROW);
    auto set_jobs = [&](int64_t first, bool in_side, bool out_side) {
        TileArgs &a = tmp.launches.back().ta;
        if (in_side) {
            a.job_in = p->d_job_in;
            a.job_swap_in = 1;
        }
        if (out_side) {
            a.job_out = p->d_job_out;
            a.job_swap_out = 1;
        }
        a.job_inv = p->d_job_inv;
        a.job_first = (int32_t)first;
        Launch &l = tmp.launches.back();
        if (in_side) l.vin_ok = l.vin_ok && p->offsets_even;
        if (out_side) l.vout_ok = l.vout_ok && p->offsets_even;
    };
    int cnt = 0;
    const KernelEntry *tab = tile_table(prec, 0, &cnt);
    for (const Cls &c : big) {
        const int64_t L = (int64_t)1 << c.lg;
        // each length class takes the kernel choices a single plan of that
        // length would: the tuning table's entry, or the defaults
        b.opts = resolve_opts(prec, L, nullptr);
        if (L <= maxrow) {
            // one transform per tile (job tables give each its own offsets): a
            // tuned single-pass kernel counts only when it holds one sequence
            const int v = b.opts.row_variant;
            const int variant = v >= 0 && v < cnt && tab[v].M == 1 ? v : -1;
            if ((st = b.tile((int)L, PAT_ROW, PAT_ROW, 1, B_IN, 0, B_OUT, 0, 1, c.count, 1, 0, 0, 0, 1, 0, 0, 0, 1,
                             TW_NONE, 0, false, false, variant)))
                return st;
            set_jobs(c.first, true, true);
            continue;
        }
        int l1, l2;
        four_step_split(prec, c.lg, &l1, &l2, b.opts.n1_log2);
        const int64_t N1 = (int64_t)1 << l1, N2 = (int64_t)1 << l2;
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(c.count, ((int64_t)1 << 30) / (L * (int64_t)cs)));
        b.need_scratch(chunk * L);
        for (int64_t b0 = 0; b0 < c.count; b0 += chunk) {
            const int64_t cb = std::min(chunk, c.count - b0);
            if ((st = b.tile((int)N1, PAT_COL, PAT_COL, b.opts.col_m, B_IN, 0, B_SCR, 0, N2, cb, 1, 0, 0, 1, N2, 0, L,
                             1, N2, TW_INNER, L, false, false)))
                return st;
            set_jobs(c.first + b0, true, false);
            if ((st = b.tile((int)N2, PAT_ROW, PAT_COL, b.opts.row_m, B_SCR, 0, B_OUT, 0, N1, cb, 1, 0, L, N2, 1, 0, 0,
                             1, N1, TW_NONE, 0, false, false)))
                return st;
            set_jobs(c.first + b0, false, true);
        }
    }
    // bounds proof of the scratch side (the job tables place the other side)
    const int64_t scratch_bytes = b.scratch_elems * (int64_t)cs;
    for (const Launch &l : tmp.launches) {
        const TileArgs &a = l.ta;
        const int64_t n_outer2 = a.n_tiles / ((int64_t)a.tiles_inner * a.n_outer1);
        for (int side = 0; side < 2; side++) {
            if ((side ? l.out_buf : l.in_buf) != B_SCR) continue;
            const int64_t last = side ? (n_outer2 - 1) * a.out_outer2 + (int64_t)(a.n_outer1 - 1) * a.out_outer +
                                            (int64_t)(a.n_inner - 1) * a.out_inner + (int64_t)(l.k->L - 1) * a.out_elem
                                      : (n_outer2 - 1) * a.in_outer2 + (int64_t)(a.n_outer1 - 1) * a.in_outer +
                                            (int64_t)(a.n_inner - 1) * a.in_inner + (int64_t)(l.k->L - 1) * a.in_elem;
            const int64_t off = side ? l.out_off : l.in_off;
            if (off < 0 || off + (last + 1) * (int64_t)cs > scratch_bytes)
                return fail(GFFT_ERR_UNSUPPORTED, "internal: batch launch leaves the scratch buffer");
        }
    }
    for (auto &l : tmp.launches) p->launches.push_back(l);
    p->scratch_bytes = (size_t)scratch_bytes;
    if (p->scratch_bytes) HIP_TRY(hipMalloc(&p->scratch, p->scratch_bytes));
    return GFFT_OK;
}

int enqueue_batch(gfft_batch_plan_s *p, const void *in, void *out, hipStream_t s)
{
    return run_launches(p->launches, p->prec, p->xcd_remap, in, out, p->scratch, GFFT_FORWARD, s);
}


// ---------------------------------------------------------------------------
// Autotuner and tuning-table persistence
// ---------------------------------------------------------------------------
int tuning_load_locked(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(GFFT_ERR_INVALID_ARG, "cannot open tuning file %s", path);
    // one entry per line: {"precision": P, "n": N, "n1_log2": A, "col_m": B, "row_m": C, "pipe_mask": D, "row_variant": E, ...}
    char line[1024];
    int loaded = 0;
    while (fgets(line, sizeof(line), f)) {
        int prec, a, b, c, d, e, du = 0;
        long long n;
        const int got = sscanf(line, " {\"precision\": %d, \"n\": %lld, \"n1_log2\": %d, \"col_m\": %d, \"row_m\": %d, "
                                     "\"pipe_mask\": %d, \"row_variant\": %d, \"dual\": %d",
                               &prec, &n, &a, &b, &c, &d, &e, &du);
        if (got >= 7) {  // files written before the dual field read as dual = 0
            PlanOpts o;
            o.n1_log2 = a;
            o.col_m = b;
            o.row_m = c;
            o.pipe_mask = d;
            o.row_variant = e;
            o.dual = got == 8 ? du : 0;
            g_tune[{prec, (int64_t)n}] = o;
            loaded++;
        }
    }
    fclose(f);
    return loaded >= 0 ? GFFT_OK : GFFT_ERR_INVALID_ARG;
}

// Candidate kernel choices for a C2C of length n.
std::vector<PlanOpts> tune_candidates(int prec, int64_t n)
{
    std::vector<PlanOpts> out;
    int cnt = 0;
    const KernelEntry *tab = prec ? tile_kernels_f64(&cnt) : tile_kernels_f32(&cnt);
    const int maxrow = max_len(prec, PAT_ROW, PAT_ROW);
    if (n <= 1) return {PlanOpts{}};
    if (n <= maxrow) {
        for (int i = 0; i < cnt; i++)
            if (tab[i].L == n && tab[i].inp == PAT_ROW && tab[i].outp == PAT_ROW) {
                PlanOpts o;
                o.row_variant = i;
                out.push_back(o);
            }
        return out;
    }
    const int lg = ilog2(n), maxcol = max_len(prec, PAT_COL, PAT_COL), maxtr = max_len(prec, PAT_ROW, PAT_COL);
    for (int l1 = 1; l1 < lg; l1++) {
        const int64_t N1 = (int64_t)1 << l1, N2 = (int64_t)1 << (lg - l1);
        if (N1 > maxcol || N2 > maxtr) continue;
        std::vector<int> cms{0}, rms{0};
        bool cpipe = false, rpipe = false;
        for (int i = 0; i < cnt; i++) {
            const KernelEntry &e = tab[i];
            if (e.L == N1 && e.inp == PAT_COL && e.outp == PAT_COL) {
                if (std::find(cms.begin(), cms.end(), e.M) == cms.end()) cms.push_back(e.M);
                cpipe = cpipe || e.persistent;
            }
            if (e.L == N2 && e.inp == PAT_ROW && e.outp == PAT_COL) {
                if (std::find(rms.begin(), rms.end(), e.M) == rms.end()) rms.push_back(e.M);
                rpipe = rpipe || e.persistent;
            }
        }
        for (int cm : cms)
            for (int rm : rms)
                for (int pm = 0; pm < 4; pm++) {
                    if (((pm & 1) && !cpipe) || ((pm & 2) && !rpipe)) continue;
                    PlanOpts o;
                    o.n1_log2 = l1;
                    o.col_m = cm;
                    o.row_m = rm;
                    o.pipe_mask = pm;
                    out.push_back(o);
                    // the overlapped form where a dual kernel pairs the two passes
                    if (pm == 3 && find_dual(prec, find_tile(prec, (int)N1, PAT_COL, PAT_COL, cm, 3),
                                             find_tile(prec, (int)N2, PAT_ROW, PAT_COL, rm, 3))) {
                        o.dual = 1;
                        out.push_back(o);
                    }
                }
    }
    return out;
}

}  // namespace
}  // namespace gfft

using namespace gfft;

extern "C" {

int gfft_autotune(int precision, int64_t n, int64_t batch, int reps, char *report, size_t len)
{
    if (precision != GFFT_F32 && precision != GFFT_F64) return fail(GFFT_ERR_INVALID_ARG, "bad precision");
    if (!is_pow2(n) || n > ((int64_t)1 << 23) || batch < 1) return fail(GFFT_ERR_UNSUPPORTED, "bad size or batch");
    if (reps < 1) reps = 10;
    int dev = 0;
    if (!device_ok(&dev)) return fail(GFFT_ERR_NO_DEVICE, "no HIP device available");
    const size_t bytes = (size_t)n * batch * (precision ? 16 : 8);
    void *in = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipMalloc(&in, bytes));
    HIP_TRY(hipMalloc(&out, bytes));
    HIP_TRY(hipMemset(in, 0x3c, bytes));
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    std::string rep = "{\"precision\": " + std::to_string(precision) + ", \"n\": " + std::to_string(n) +
                      ", \"batch\": " + std::to_string(batch) + ", \"candidates\": [\n";
    float best_ms = 1e30f;
    PlanOpts best;
    int st = GFFT_OK, tried = 0;
    for (const PlanOpts &o : tune_candidates(precision, n)) {
        std::unique_ptr<gfft_plan_s> p(new gfft_plan_s());
        p->d = gfft_desc{};
        p->d.kind = GFFT_C2C;
        p->d.precision = precision;
        p->d.n = n;
        p->d.batch = batch;
        p->prec = precision;
        p->device = dev;
        if (hipDeviceGetAttribute(&p->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            p->num_cus = 256;
        if (build_plan(p.get(), false, &o) || validate(p.get())) continue;  // candidate not plannable
        if (p->scratch_bytes && hipMalloc(&p->scratch, p->scratch_bytes) != hipSuccess) continue;
        float ms = 0;
        int r = run(p.get(), in, out, GFFT_FORWARD, s);  // warm-up
        r = r ? r : run(p.get(), in, out, GFFT_FORWARD, s);
        if (!r && hipEventRecord(e0, s) == hipSuccess) {
            for (int i = 0; i < reps && !r; i++) r = run(p.get(), in, out, GFFT_FORWARD, s);
            if (!r && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                ms /= reps;
            else
                r = GFFT_ERR_HIP;
        }
        if (p->scratch) (void)hipFree(p->scratch);
        p->scratch = nullptr;
        if (r) {
            st = r;
            break;
        }
        char line[256];
        snprintf(line, sizeof(line),
                 "  {\"n1_log2\": %d, \"col_m\": %d, \"row_m\": %d, \"pipe_mask\": %d, \"row_variant\": %d, "
                 "\"dual\": %d, \"ms\": %.5f}",
                 o.n1_log2, o.col_m, o.row_m, o.pipe_mask, o.row_variant, o.dual, ms);
        rep += (tried ? ",\n" : "") + std::string(line);
        tried++;
        if (ms < best_ms) {
            best_ms = ms;
            best = o;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    (void)hipFree(in);
    (void)hipFree(out);
    if (st) return st;
    if (!tried) return fail(GFFT_ERR_UNSUPPORTED, "no candidate could be planned");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_loaded = true;
        g_tune[{precision, n}] = best;
    }
    char tail[256];
    snprintf(tail, sizeof(tail),
             "\n], \"best\": {\"n1_log2\": %d, \"col_m\": %d, \"row_m\": %d, \"pipe_mask\": %d, \"row_variant\": %d, "
             "\"dual\": %d, \"ms\": %.5f}}",
             best.n1_log2, best.col_m, best.row_m, best.pipe_mask, best.row_variant, best.dual, best_ms);
    rep += tail;
    if (report && len) snprintf(report, len, "%s", rep.c_str());
    return GFFT_OK;
}

int gfft_tuning_save(const char *path)
{
    if (!path) return fail(GFFT_ERR_INVALID_ARG, "null path");
    FILE *f = fopen(path, "w");
    if (!f) return fail(GFFT_ERR_INVALID_ARG, "cannot write %s", path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    fprintf(f, "{\"version\": 1, \"entries\": [\n");
    size_t i = 0;
    for (const auto &kv : g_tune) {
        const PlanOpts &o = kv.second;
        fprintf(f,
                " {\"precision\": %d, \"n\": %lld, \"n1_log2\": %d, \"col_m\": %d, \"row_m\": %d, \"pipe_mask\": %d, "
                "\"row_variant\": %d, \"dual\": %d}%s\n",
                kv.first.first, (long long)kv.first.second, o.n1_log2, o.col_m, o.row_m, o.pipe_mask, o.row_variant,
                o.dual, ++i < g_tune.size() ? "," : "");
    }
    fprintf(f, "]}\n");
    fclose(f);
    return GFFT_OK;
}

int gfft_tuning_load(const char *path)
{
    if (!path) return fail(GFFT_ERR_INVALID_ARG, "null path");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_loaded = true;
    return tuning_load_locked(path);
}

int gfft_tuning_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    g_tune_loaded = true;
    return GFFT_OK;
}

}  // extern "C"

namespace gfft {
namespace {
}  // namespace
}  // namespace gfft

// ===========================================================================
// C ABI
// ===========================================================================
using namespace gfft;

extern "C" {

int gfft_version(void) { return GFFT_VERSION; }

const char *gfft_error_string(int status)
{
    switch (status) {
    case GFFT_OK: return "success";
    case GFFT_ERR_INVALID_ARG: return "invalid argument";
    case GFFT_ERR_UNSUPPORTED: return "unsupported size or layout";
    case GFFT_ERR_HIP: return "HIP runtime error";
    case GFFT_ERR_ALLOC: return "allocation failed";
    case GFFT_ERR_NO_DEVICE: return "no usable GPU device";
    default: return "unknown status";
    }
}

const char *gfft_last_error(void) { return g_last_error.c_str(); }

int gfft_is_device_pointer(const void *ptr)
{
    if (!ptr) return 0;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

static int check_desc(const gfft_desc &d)
{
    if (d.precision != GFFT_F32 && d.precision != GFFT_F64 && d.precision != GFFT_F16 && d.precision != GFFT_BF16)
        return fail(GFFT_ERR_INVALID_ARG, "precision must be GFFT_F32, GFFT_F64, GFFT_F16 or GFFT_BF16");
    if (d.kind < GFFT_C2C || d.kind > GFFT_C2R_2D) return fail(GFFT_ERR_INVALID_ARG, "unknown kind %d", d.kind);
    if (d.batch < 1) return fail(GFFT_ERR_INVALID_ARG, "batch must be >= 1");
    // DIT (the real-FFT unpack alone) also takes any n divisible by 4, as
    // genFFT's dynamic FFTDIT<-1> does (generic/fft_dit_impl_generic.inl:117-153)
    const bool dit_any = d.kind == GFFT_DIT && d.n > 0 && d.n % 4 == 0;
    if ((!is_pow2(d.n) && !dit_any) || d.n > ((int64_t)1 << 23))
        return fail(GFFT_ERR_UNSUPPORTED, "size %lld is not a power of two in 1..2^23%s", (long long)d.n,
                    d.kind == GFFT_DIT ? " or a multiple of 4" : "");
    const bool two = d.kind == GFFT_C2C_2D || d.kind == GFFT_R2C_2D || d.kind == GFFT_COLS || d.kind == GFFT_C2R_2D;
    if (two && d.kind != GFFT_COLS && (!is_pow2(d.n2) || d.n2 > ((int64_t)1 << 23)))
        return fail(GFFT_ERR_UNSUPPORTED, "height %lld is not a power of two", (long long)d.n2);
    if (d.kind == GFFT_COLS && d.n2 < 1) return fail(GFFT_ERR_INVALID_ARG, "column count must be >= 1");
    if (d.in_stride < 0 || d.out_stride < 0 || d.in_dist < 0 || d.out_dist < 0)
        return fail(GFFT_ERR_INVALID_ARG, "negative stride or distance");
    if (d.reserved != 0) return fail(GFFT_ERR_INVALID_ARG, "reserved field must be 0");
    const int64_t width = d.kind == GFFT_COLS ? d.n2 : d.n;
    const int64_t in_width = d.kind == GFFT_C2R_2D ? d.n / 2 + 1 : width;  // C2R_2D reads n/2+1 columns
    if (two && ((d.in_stride && d.in_stride < in_width) || (d.out_stride && d.out_stride < width)))
        return fail(GFFT_ERR_INVALID_ARG, "row pitch smaller than the row length");
    if (d.precision == GFFT_F16 || d.precision == GFFT_BF16) {
        if (d.kind != GFFT_C2C && d.kind != GFFT_C2C_2D && d.kind != GFFT_COLS)
            return fail(GFFT_ERR_UNSUPPORTED, "16-bit storage supports C2C, C2C_2D and COLS plans");
        if (d.n < 2) return fail(GFFT_ERR_UNSUPPORTED, "16-bit storage needs a transform length >= 2");
    }
    return GFFT_OK;
}

static void set_precision(gfft_plan_s *p, int precision)
{
    p->prec = precision == GFFT_F64 ? 1 : 0;
    p->store = precision == GFFT_F16 ? 1 : precision == GFFT_BF16 ? 2 : 0;
}

int gfft_plan_explain(const gfft_desc *desc, char *buf, size_t len)
{
    if (!desc || !buf || !len) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    int st = check_desc(*desc);
    if (st) return st;
    std::unique_ptr<gfft_plan_s> p(new gfft_plan_s());
    p->d = *desc;
    set_precision(p.get(), desc->precision);
    if ((st = build_plan(p.get(), true))) return st;
    if ((st = validate(p.get()))) return st;
    if ((st = check_noalloc_fit(p.get()))) return st;
    return describe(p.get(), buf, len);
}

int gfft_plan_create(gfft_plan *plan, const gfft_desc *desc)
{
    if (!plan || !desc) return fail(GFFT_ERR_INVALID_ARG, "null plan or descriptor");
    *plan = nullptr;
    const gfft_desc &d = *desc;
    int st = check_desc(d);
    if (st) return st;
    int dev = 0;
    if (!device_ok(&dev)) return fail(GFFT_ERR_NO_DEVICE, "no HIP device available");
    std::unique_ptr<gfft_plan_s> p(new gfft_plan_s());
    p->d = d;
    set_precision(p.get(), d.precision);
    p->device = dev;
    p->xcd_remap = (int)env_int("GFFT_XCD_REMAP", 1);
    if (hipDeviceGetAttribute(&p->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || p->num_cus < 1)
        p->num_cus = 256;
    if ((st = build_plan(p.get(), false))) return st;
    if ((st = validate(p.get()))) return st;
    if (d.flags & GFFT_FLAG_NO_WORKSPACE) {
        // set-up for gfft_execute_noalloc on this device: resolve the static
        // workspace's address now, so the calls themselves touch no runtime table
        void *scr = nullptr;
        if ((st = check_noalloc_fit(p.get()))) return st;
        if (p->scratch_bytes && (st = noalloc_scratch_ptr(dev, &scr))) return st;
    } else if (p->scratch_bytes) {
        hipError_t e = hipMalloc(&p->scratch, p->scratch_bytes);
        if (e != hipSuccess)
            return fail(GFFT_ERR_ALLOC, "scratch of %zu bytes: %s", p->scratch_bytes, hipGetErrorString(e));
    }
    *plan = p.release();
    return GFFT_OK;
}

int gfft_plan_destroy(gfft_plan plan)
{
    if (!plan) return GFFT_OK;
    if (plan->scratch) (void)hipFree(plan->scratch);
    if (plan->hin) (void)hipFree(plan->hin);
    if (plan->hout) (void)hipFree(plan->hout);
    if (plan->pin_in) (void)hipHostFree(plan->pin_in);
    if (plan->pin_out) (void)hipHostFree(plan->pin_out);
    if (plan->host_stream) (void)hipStreamDestroy(plan->host_stream);
    if (plan->perm) (void)hipFree(plan->perm);
    if (plan->stage) (void)hipFree(plan->stage);
    delete plan;
    return GFFT_OK;
}

int gfft_execute(gfft_plan plan, const void *in, void *out, int direction, void *stream)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    if (!in || !out) return fail(GFFT_ERR_INVALID_ARG, "null buffer");
    if (direction != GFFT_FORWARD && direction != GFFT_INVERSE)
        return fail(GFFT_ERR_INVALID_ARG, "direction must be GFFT_FORWARD or GFFT_INVERSE");
    if (!(plan->dir_mask & (1 << direction)))
        return fail(GFFT_ERR_INVALID_ARG, "this plan kind supports only the %s direction",
                    plan->dir_mask == 1 ? "forward" : "inverse");
    if (int st = on_plan_device(plan->device)) return st;
    return run(plan, in, out, direction, (hipStream_t)stream);
}

int gfft_execute_noalloc(gfft_plan plan, const void *in, void *out, int direction, void *stream)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    if (!in || !out) return fail(GFFT_ERR_INVALID_ARG, "null buffer");
    if (direction != GFFT_FORWARD && direction != GFFT_INVERSE)
        return fail(GFFT_ERR_INVALID_ARG, "direction must be GFFT_FORWARD or GFFT_INVERSE");
    if (!(plan->dir_mask & (1 << direction)))
        return fail(GFFT_ERR_INVALID_ARG, "this plan kind supports only the %s direction",
                    plan->dir_mask == 1 ? "forward" : "inverse");
    if (int st = on_plan_device(plan->device)) return st;
    if (plan->scratch_bytes > (size_t)GFFT_NOALLOC_SCRATCH_BYTES)
        return fail(GFFT_ERR_UNSUPPORTED,
                    "this plan needs %zu bytes of workspace, more than the %lld bytes of the static workspace "
                    "(create it with GFFT_FLAG_NO_WORKSPACE to chunk its batch)",
                    plan->scratch_bytes, (long long)GFFT_NOALLOC_SCRATCH_BYTES);
    if (needs_stage(plan, in, out))
        return fail(GFFT_ERR_UNSUPPORTED, "the buffers overlap other than exactly in place, which needs a staging "
                                          "copy this entry point cannot allocate");
    void *scr = nullptr;
    if (plan->scratch_bytes)
        if (int st = noalloc_scratch_ptr(plan->device, &scr)) return st;
    return run_launches(plan->launches, plan->prec, plan->xcd_remap, in, out, scr, direction, (hipStream_t)stream);
}

int gfft_execute_scaled(gfft_plan plan, const void *in, void *out, int direction, float *scales, int scale_mode,
                        void *stream)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    if (!in || !out || !scales) return fail(GFFT_ERR_INVALID_ARG, "null buffer");
    if (direction != GFFT_FORWARD && direction != GFFT_INVERSE)
        return fail(GFFT_ERR_INVALID_ARG, "direction must be GFFT_FORWARD or GFFT_INVERSE");
    if (!plan->store) return fail(GFFT_ERR_UNSUPPORTED, "per-transform scales need a GFFT_F16 / GFFT_BF16 plan");
    if (scale_mode != GFFT_SCALE_GIVEN && scale_mode != GFFT_SCALE_AUTO)
        return fail(GFFT_ERR_INVALID_ARG, "scale_mode must be GFFT_SCALE_GIVEN or GFFT_SCALE_AUTO");
    if (int st = on_plan_device(plan->device)) return st;
    hipStream_t s = (hipStream_t)stream;
    if (scale_mode == GFFT_SCALE_AUTO) {
        AmaxArgs a{};
        a.in = in;
        a.bits = reinterpret_cast<uint32_t *>(scales);
        a.store = plan->store;
        a.batch = plan->d.batch;
        a.rows = plan->amax_rows;
        a.width = plan->amax_width;
        a.pitch = plan->amax_pitch;
        a.dist = plan->scale_div;
        a.gain = (float)plan->amax_gain;
        // largest bound a stored value may reach: 2^15 keeps FP16 (max 65504)
        // finite with room for rounding; BF16 shares FP32's exponent range
        a.limit = plan->store == 1 ? 32768.0f : 0x1p120f;
        HIP_TRY(launch_amax_scale(a, s));
    }
    return run(plan, in, out, direction, s, nullptr, scales);
}

int gfft_execute_scrambled(gfft_plan plan, const void *in, void *out, int direction, void *stream)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    if (!in || !out) return fail(GFFT_ERR_INVALID_ARG, "null buffer");
    if (direction != GFFT_FORWARD && direction != GFFT_INVERSE)
        return fail(GFFT_ERR_INVALID_ARG, "direction must be GFFT_FORWARD or GFFT_INVERSE");
    const gfft_desc &d = plan->d;
    if ((d.kind != GFFT_C2C && d.kind != GFFT_COLS) || plan->store)
        return fail(GFFT_ERR_UNSUPPORTED, "scrambled input needs a complex64 / complex128 C2C or COLS plan");
    if (int st = on_plan_device(plan->device)) return st;
    if (plan->scratch_bytes && !plan->scratch)
        return fail(GFFT_ERR_INVALID_ARG, "plan created with GFFT_FLAG_NO_WORKSPACE holds no workspace: "
                                          "run it with gfft_execute_noalloc");
    BitrevArgs a{};
    a.n = d.n;
    a.lg = ilog2(d.n);
    a.batch = d.batch;
    if (d.kind == GFFT_C2C) {
        a.width = 1;
        a.stride = 1;
        a.dist = dflt(d.in_dist, d.n);
    } else {
        a.width = d.n2;
        a.stride = dflt(d.in_stride, d.n2);
        a.dist = dflt(d.in_dist, d.n * a.stride);
    }
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(plan->order.mu);
    if (!plan->perm) {
        hipError_t e = hipMalloc(&plan->perm, plan->in_bytes);
        if (e != hipSuccess) return fail(GFFT_ERR_ALLOC, "scramble buffer of %zu bytes: %s", plan->in_bytes,
                                         hipGetErrorString(e));
    }
    HIP_TRY(plan->order.before(s));
    a.in = in;
    a.out = plan->perm;
    HIP_TRY(launch_bitrev(plan->prec, a, s));
    int st = run_launches(plan->launches, plan->prec, plan->xcd_remap, plan->perm, out, plan->scratch, direction, s,
                          nullptr, nullptr, plan->scale_div);
    if (st) return st;
    HIP_TRY(plan->order.after(s));
    return GFFT_OK;
}

int gfft_execute_host(gfft_plan plan, const void *in, void *out, int direction)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    if (!in || !out) return fail(GFFT_ERR_INVALID_ARG, "null buffer");
    if (int st = on_plan_device(plan->device)) return st;
    std::lock_guard<std::mutex> lk(plan->host_mu);
    // Calls up to 1 MiB a side: the launches read and write plan-owned pinned
    // (coherent) host buffers in place, on the plan's own stream, with one
    // synchronisation and no copy commands.  Measured on MI355X, c64 forward +
    // inverse pairs: n = 2^14 98 -> 71 us, 2^17 270 -> 250 us against DMA
    // copies through the same buffers; from 2 MiB on the runtime's pageable
    // copies were as fast or faster (2^20: 1.57 against 1.68 ms in place).
    constexpr size_t kPinnedMax = (size_t)1 << 20;
    if (plan->in_bytes <= kPinnedMax && plan->out_bytes <= kPinnedMax) {
        if (!plan->pin_in || !plan->pin_out || !plan->host_stream) {
            void *pi = nullptr, *po = nullptr;
            hipStream_t hs = nullptr;
            bool ok = hipHostMalloc(&pi, plan->in_bytes, hipHostMallocCoherent) == hipSuccess &&
                      hipHostMalloc(&po, plan->out_bytes, hipHostMallocCoherent) == hipSuccess &&
                      hipStreamCreateWithFlags(&hs, hipStreamNonBlocking) == hipSuccess;
            if (!ok) {
                if (pi) (void)hipHostFree(pi);
                if (po) (void)hipHostFree(po);
                return fail(GFFT_ERR_ALLOC, "pinned staging of %zu + %zu bytes", plan->in_bytes, plan->out_bytes);
            }
            if (plan->pin_in) (void)hipHostFree(plan->pin_in);
            if (plan->pin_out) (void)hipHostFree(plan->pin_out);
            if (plan->host_stream) (void)hipStreamDestroy(plan->host_stream);
            plan->pin_in = pi;
            plan->pin_out = po;
            plan->host_stream = hs;
        }
        const hipStream_t hs = plan->host_stream;
        void *di = nullptr, *dout = nullptr;
        hipError_t e = hipHostGetDevicePointer(&di, plan->pin_in, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&dout, plan->pin_out, 0);
        if (e != hipSuccess) return fail(GFFT_ERR_HIP, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
        memcpy(plan->pin_in, in, plan->in_bytes);
        if (plan->out_has_gaps) memcpy(plan->pin_out, out, plan->out_bytes);  // the bytes no launch writes
        int st = gfft_execute(plan, di, dout, direction, hs);
        // wait for whatever was enqueued, also on failure: the next call reuses the buffers
        e = hipStreamSynchronize(hs);
        if (st) return st;
        if (e != hipSuccess) return fail(GFFT_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
        memcpy(out, plan->pin_out, plan->out_bytes);
        return GFFT_OK;
    }
    // larger calls: the runtime's copies through device staging buffers
    if (!plan->hin || !plan->hout) {
        // both staging buffers or neither: a failed second allocation rolls back the first
        void *hin = nullptr, *hout = nullptr;
        if (hipMalloc(&hin, plan->in_bytes) != hipSuccess)
            return fail(GFFT_ERR_ALLOC, "device staging of %zu bytes", plan->in_bytes);
        if (hipMalloc(&hout, plan->out_bytes) != hipSuccess) {
            (void)hipFree(hin);
            return fail(GFFT_ERR_ALLOC, "device staging of %zu bytes", plan->out_bytes);
        }
        if (plan->hin) (void)hipFree(plan->hin);
        if (plan->hout) (void)hipFree(plan->hout);
        plan->hin = hin;
        plan->hout = hout;
    }
    HIP_TRY(hipMemcpy(plan->hin, in, plan->in_bytes, hipMemcpyHostToDevice));
    // keep the bytes the transform does not write (pitch gaps) intact
    if (plan->out_has_gaps) HIP_TRY(hipMemcpy(plan->hout, out, plan->out_bytes, hipMemcpyHostToDevice));
    int st = gfft_execute(plan, plan->hin, plan->hout, direction, nullptr);
    if (st) return st;
    HIP_TRY(hipMemcpy(out, plan->hout, plan->out_bytes, hipMemcpyDeviceToHost));
    return GFFT_OK;
}

int gfft_plan_profile(gfft_plan plan, const void *in, void *out, int direction, void *stream, float *ms,
                      int max_launches, int *count)
{
    if (!plan || !in || !out || !count) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    if ((direction != GFFT_FORWARD && direction != GFFT_INVERSE) || !(plan->dir_mask & (1 << direction)))
        return fail(GFFT_ERR_INVALID_ARG, "direction %d is not one this plan runs", direction);
    if (int st = on_plan_device(plan->device)) return st;
    const int n = (int)plan->launches.size();
    *count = n;
    std::vector<hipEvent_t> ev(2 * n);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    int st = run(plan, in, out, direction, (hipStream_t)stream, ev.data());
    if (!st) {
        hipError_t e = hipEventSynchronize(ev[2 * n - 1]);
        if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "hipEventSynchronize: %s", hipGetErrorString(e));
    }
    for (int i = 0; !st && ms && i < n && i < max_launches; i++) {
        hipError_t e = hipEventElapsedTime(&ms[i], ev[2 * i], ev[2 * i + 1]);
        if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "hipEventElapsedTime: %s", hipGetErrorString(e));
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return st;
}

int gfft_plan_profile_launch(gfft_plan plan, const void *in, void *out, int direction, void *stream, int launch,
                             int reps, float *ms)
{
    if (!plan || !in || !out || !ms) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    if (launch < 0 || launch >= (int)plan->launches.size())
        return fail(GFFT_ERR_INVALID_ARG, "launch %d out of range (plan has %zu)", launch, plan->launches.size());
    if (reps < 1 || reps > 10000) return fail(GFFT_ERR_INVALID_ARG, "reps %d out of range 1 .. 10000", reps);
    if ((direction != GFFT_FORWARD && direction != GFFT_INVERSE) || !(plan->dir_mask & (1 << direction)))
        return fail(GFFT_ERR_INVALID_ARG, "direction %d is not one this plan runs", direction);
    const Launch &l = plan->launches[launch];
    if (l.type == LT_MIX || l.type == LT_MIX_RESET)
        return fail(GFFT_ERR_UNSUPPORTED, "a work-queue launch runs once per reset");
    if (plan->scratch_bytes && !plan->scratch)
        return fail(GFFT_ERR_INVALID_ARG, "plan created with GFFT_FLAG_NO_WORKSPACE holds no workspace");
    if (int st = on_plan_device(plan->device)) return st;
    const hipStream_t s = (hipStream_t)stream;
    const std::vector<Launch> rep((size_t)reps, l);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    int st = GFFT_OK;
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
    if (!st) {
        std::lock_guard<std::mutex> lk(plan->order.mu);  // the workspace is shared with other calls
        e = plan->order.before(s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "%s", hipGetErrorString(e));
        if (!st) st = run_launches(rep, plan->prec, plan->xcd_remap, in, out, plan->scratch, direction, s);
        if (!st) {
            e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = plan->order.after(s);
            if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "%s", hipGetErrorString(e));
        }
    }
    if (!st) {
        float t = 0.f;
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) st = fail(GFFT_ERR_HIP, "%s", hipGetErrorString(e));
        else *ms = t / (float)reps;
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return st;
}

int gfft_plan_workspace(gfft_plan plan, size_t *bytes)
{
    if (!plan || !bytes) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    *bytes = plan->scratch_bytes;
    return GFFT_OK;
}

int gfft_plan_extents(gfft_plan plan, size_t *in_bytes, size_t *out_bytes)
{
    if (!plan || !in_bytes || !out_bytes) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    *in_bytes = plan->in_bytes;
    *out_bytes = plan->out_bytes;
    return GFFT_OK;
}

int gfft_plan_describe(gfft_plan plan, char *buf, size_t len)
{
    if (!plan || !buf || !len) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    return describe(plan, buf, len);
}

int gfft_separate_2x_real(int precision, void *out1, void *out2, const void *in, int64_t n, int64_t batch,
                          void *stream)
{
    if (!out1 || !out2 || !in || n < 1 || batch < 1) return fail(GFFT_ERR_INVALID_ARG, "bad argument");
    SeparateArgs a{in, out1, out2, n, batch, n};
    HIP_TRY(launch_separate(precision, a, (hipStream_t)stream));
    return GFFT_OK;
}

int gfft_interleave_real(int precision, void *out, const void *in1, const void *in2, int64_t n, int64_t batch,
                         void *stream)
{
    if (!out || !in1 || !in2 || n < 1 || batch < 1) return fail(GFFT_ERR_INVALID_ARG, "bad argument");
    InterleaveArgs a{in1, in2, out, n, batch, n, n, n};
    HIP_TRY(launch_interleave(precision, a, (hipStream_t)stream));
    return GFFT_OK;
}

int gfft_interleave_real_2d(int precision, void *out, int64_t out_stride, const void *in1, int64_t in_stride1,
                            const void *in2, int64_t in_stride2, int64_t width, int64_t height, void *stream)
{
    if (!out || !in1 || !in2 || width < 1 || height < 1) return fail(GFFT_ERR_INVALID_ARG, "bad argument");
    if (precision != GFFT_F32 && precision != GFFT_F64) return fail(GFFT_ERR_INVALID_ARG, "bad precision");
    if (out_stride < width || in_stride1 < width || in_stride2 < width)
        return fail(GFFT_ERR_INVALID_ARG, "row pitch smaller than the width");
    InterleaveArgs a{in1, in2, out, width, height, in_stride1, in_stride2, out_stride};
    HIP_TRY(launch_interleave(precision, a, (hipStream_t)stream));
    return GFFT_OK;
}

}  // extern "C"

extern "C" {

int gfft_batch_plan_create(gfft_batch_plan *plan, int precision, int64_t count, const int64_t *sizes,
                           const int64_t *in_offsets, const int64_t *out_offsets, const int32_t *directions)
{
    if (!plan || !sizes || !in_offsets || !out_offsets || !directions || count < 1)
        return fail(GFFT_ERR_INVALID_ARG, "null argument or empty batch");
    *plan = nullptr;
    if (precision != GFFT_F32 && precision != GFFT_F64) return fail(GFFT_ERR_INVALID_ARG, "bad precision");
    for (int64_t j = 0; j < count; j++) {
        if (!is_pow2(sizes[j]) || sizes[j] > ((int64_t)1 << 23))
            return fail(GFFT_ERR_UNSUPPORTED, "job %lld: size %lld is not a power of two in 1..2^23", (long long)j,
                        (long long)sizes[j]);
        if (in_offsets[j] < 0 || out_offsets[j] < 0)
            return fail(GFFT_ERR_INVALID_ARG, "job %lld: negative offset", (long long)j);
        if (directions[j] != GFFT_FORWARD && directions[j] != GFFT_INVERSE)
            return fail(GFFT_ERR_INVALID_ARG, "job %lld: bad direction", (long long)j);
    }
    int dev = 0;
    if (!device_ok(&dev)) return fail(GFFT_ERR_NO_DEVICE, "no HIP device available");
    std::unique_ptr<gfft_batch_plan_s> p(new gfft_batch_plan_s());
    p->prec = precision;
    p->device = dev;
    p->count = count;
    p->sizes.assign(sizes, sizes + count);
    p->in_off.assign(in_offsets, in_offsets + count);
    p->out_off.assign(out_offsets, out_offsets + count);
    p->dirs.assign(directions, directions + count);
    p->xcd_remap = (int)env_int("GFFT_XCD_REMAP", 1);
    if (hipDeviceGetAttribute(&p->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || p->num_cus < 1)
        p->num_cus = 256;
    int st = build_batch(p.get());
    if (st) {
        gfft_batch_plan_destroy(p.release());
        return st;
    }
    *plan = p.release();
    return GFFT_OK;
}

int gfft_batch_plan_destroy(gfft_batch_plan plan)
{
    if (!plan) return GFFT_OK;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->cap_stream) (void)hipStreamDestroy(plan->cap_stream);
    for (auto &kv : plan->per_job) gfft_plan_destroy(kv.second);
    for (void *ptr : {(void *)plan->d_job_in, (void *)plan->d_job_out, (void *)plan->d_job_inv,
                      (void *)plan->d_tasks, (void *)plan->d_counter, plan->scratch})
        if (ptr) (void)hipFree(ptr);
    delete plan;
    return GFFT_OK;
}

int gfft_batch_execute(gfft_batch_plan plan, const void *in, void *out, int mode, void *stream)
{
    if (!plan || !in || !out) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    if (int st = on_plan_device(plan->device)) return st;
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(plan->mu);
    if (mode == GFFT_BATCH_EAGER) {
        std::lock_guard<std::mutex> lo(plan->order.mu);
        HIP_TRY(plan->order.before(s));
        int st = enqueue_batch(plan, in, out, s);
        if (st) return st;
        HIP_TRY(plan->order.after(s));
        return GFFT_OK;
    }
    if (mode == GFFT_BATCH_PER_JOB) {
        // the plain path: one single-transform plan launch per job
        for (int64_t j = 0; j < plan->count; j++) {
            gfft_plan &pj = plan->per_job[plan->sizes[j]];
            if (!pj) {
                gfft_desc d{};
                d.kind = GFFT_C2C;
                d.precision = plan->prec;
                d.n = plan->sizes[j];
                d.batch = 1;
                int st = gfft_plan_create(&pj, &d);
                if (st) return st;
            }
            const size_t cs = plan->prec ? 16 : 8;
            int st = gfft_execute(pj, (const char *)in + plan->in_off[j] * cs, (char *)out + plan->out_off[j] * cs,
                                  plan->dirs[j], stream);
            if (st) return st;
        }
        return GFFT_OK;
    }
    if (mode != GFFT_BATCH_GRAPH) return fail(GFFT_ERR_INVALID_ARG, "unknown batch mode %d", mode);
    std::lock_guard<std::mutex> lo(plan->order.mu);
    if (!plan->exec || plan->g_in != in || plan->g_out != out) {
        if (plan->exec) {
            // the previous graph may still run on some stream: let it finish
            if (plan->order.used) HIP_TRY(hipEventSynchronize(plan->order.done));
            (void)hipGraphExecDestroy(plan->exec);
            (void)hipGraphDestroy(plan->graph);
            plan->exec = nullptr;
            plan->graph = nullptr;
        }
        if (!plan->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue_batch(plan, in, out, plan->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(plan->cap_stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return fail(GFFT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        plan->graph = g;
        HIP_TRY(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
        plan->g_in = in;
        plan->g_out = out;
    }
    HIP_TRY(plan->order.before(s));
    HIP_TRY(hipGraphLaunch(plan->exec, s));
    HIP_TRY(plan->order.after(s));
    return GFFT_OK;
}

int gfft_batch_plan_describe(gfft_batch_plan plan, char *buf, size_t len)
{
    if (!plan || !buf || !len) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    char head[256];
    snprintf(head, sizeof(head), "batch jobs=%lld scratch_bytes=%zu offsets_even=%d\n", (long long)plan->count,
             plan->scratch_bytes, (int)plan->offsets_even);
    return describe_launches(plan->launches, head, buf, len);
}

}  // extern "C"

// ===========================================================================
// Fused spectral filter (kern_filter.hip)
// ===========================================================================
struct gfft_filter_plan_s {
    int prec = 0, device = 0;
    int64_t n = 0, batch = 0;
    void *tw = nullptr;        // W_n^k, k < n/2
    const void *pass_tw = nullptr;      // shared device tables (not owned): the forward FFT's passes,
    const void *pass_tw_inv = nullptr;  // and the inverse's (its own radix order)
    void *partial = nullptr;   // batch rows of n/2 + 1 complex, filter_partial_pitch apart
    StreamOrder order;         // the partial buffer is shared by backward calls
    ~gfft_filter_plan_s()
    {
        if (tw) (void)hipFree(tw);
        if (partial) (void)hipFree(partial);
    }
};

extern "C" {

int gfft_filter_plan_create(gfft_filter_plan *plan, int precision, int64_t n, int64_t batch)
{
    if (!plan) return fail(GFFT_ERR_INVALID_ARG, "null plan");
    *plan = nullptr;
    if (precision != GFFT_F32 && precision != GFFT_F64) return fail(GFFT_ERR_INVALID_ARG, "bad precision");
    if (batch < 1 || batch > INT32_MAX) return fail(GFFT_ERR_INVALID_ARG, "bad batch");
    const int64_t L = n / 2;
    KernelEntry k{};
    k.prec = precision;
    k.L = (int)L;
    int nt = 0;
    if (!is_pow2(n) || n < 64 || (k.nrad = filter_shape(precision, (int)L, &nt, k.rad, 0)) == 0)
        return fail(GFFT_ERR_UNSUPPORTED, "fused filter length %lld not supported (64 .. %d)", (long long)n,
                    precision ? 8192 : 16384);
    KernelEntry ki = k;
    ki.nrad = filter_shape(precision, (int)L, &nt, ki.rad, 1);
    int dev = 0;
    if (!device_ok(&dev)) return fail(GFFT_ERR_NO_DEVICE, "no HIP device available");
    std::unique_ptr<gfft_filter_plan_s> p(new gfft_filter_plan_s());
    p->prec = precision;
    p->device = dev;
    p->n = n;
    p->batch = batch;
    int st = get_pass_twiddles(dev, k, &p->pass_tw, false);
    if (st) return st;
    if ((st = get_pass_twiddles(dev, ki, &p->pass_tw_inv, false))) return st;
    std::vector<long double> re(L), im(L);
    for (int64_t m = 0; m < L; m++) twiddle_ld(m, n, &re[m], &im[m]);
    if (precision) {
        std::vector<double> h(2 * L);
        for (int64_t m = 0; m < L; m++) h[2 * m] = (double)re[m], h[2 * m + 1] = (double)im[m];
        st = upload(&p->tw, h.data(), h.size() * sizeof(double));
    } else {
        std::vector<float> h(2 * L);
        for (int64_t m = 0; m < L; m++) h[2 * m] = (float)re[m], h[2 * m + 1] = (float)im[m];
        st = upload(&p->tw, h.data(), h.size() * sizeof(float));
    }
    if (st) return st;
    const size_t cs = precision ? 16 : 8;
    HIP_TRY(hipMalloc(&p->partial, (size_t)batch * filter_partial_pitch(L + 1, precision) * cs));
    *plan = p.release();
    return GFFT_OK;
}

int gfft_filter_plan_destroy(gfft_filter_plan plan)
{
    delete plan;  // frees the plan's device buffers
    return GFFT_OK;
}

int gfft_filter_forward(gfft_filter_plan plan, const void *x, const void *H, void *y, int conj_h, void *stream)
{
    if (!plan || !x || !H || !y) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    if (int st = on_plan_device(plan->device)) return st;
    FilterArgs a{};
    a.x = x;
    a.H = H;
    a.y = y;
    a.tw = plan->tw;
    a.pass_tw = plan->pass_tw;
    a.pass_tw_inv = plan->pass_tw_inv;
    a.batch = plan->batch;
    a.bins = plan->n / 2 + 1;
    a.conj_h = conj_h ? 1 : 0;
    HIP_TRY(launch_filter(plan->prec, 0, a, (hipStream_t)stream));
    return GFFT_OK;
}

int gfft_filter_backward(gfft_filter_plan plan, const void *x, const void *H, const void *gy, void *gx, void *gH,
                         void *stream)
{
    if (!plan || !x || !H || !gy || !gx || !gH) return fail(GFFT_ERR_INVALID_ARG, "null argument");
    if (int st = on_plan_device(plan->device)) return st;
    FilterArgs a{};
    a.x = x;
    a.gy = gy;
    a.H = H;
    a.y = gx;
    a.partial = plan->partial;
    a.tw = plan->tw;
    a.pass_tw = plan->pass_tw;
    a.pass_tw_inv = plan->pass_tw_inv;
    a.batch = plan->batch;
    a.bins = plan->n / 2 + 1;
    a.ppitch = filter_partial_pitch(a.bins, plan->prec);
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(plan->order.mu);
    HIP_TRY(plan->order.before(s));
    HIP_TRY(launch_filter(plan->prec, 1, a, s));
    a.y = gH;
    const hipEvent_t stop = plan->order.stop_event(s);
    HIP_TRY(launch_filter(plan->prec, 2, a, s, stop));
    if (stop) plan->order.mark(s);
    else HIP_TRY(plan->order.after(s));
    return GFFT_OK;
}

}  // extern "C"
