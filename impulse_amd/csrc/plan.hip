// Planner: turns an n-D c2c / r2c / c2r request into a list of GPU passes.
//
// Composition over axes follows pocketfft_hdronly.h:3272-3390:
//   c2c : one job per entry of `axes`, in order; the first reads the input, the
//         rest work in place on the output; fct is applied once (first job).
//   r2c : r2c along axes.back() (input -> output, fct), then c2c in place on
//         the output over the remaining axes.
//   c2r : c2c over all axes but the last into a contiguous temporary, then
//         c2r along axes.back() (temporary -> output, fct).
// Each 1-D job picks one algorithm (see build_job):
//   length <= 8192 and 13-smooth  : one LDS pass (real transforms fused)
//   13-smooth, longer             : four-step, two LDS passes through a workspace
//   otherwise                     : Bluestein chirp-z with a 7-smooth length
//                                   n2 >= 2L-1 (one LDS pass, or four passes)
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "butterflies.h"
#include "engine.h"

namespace ifft {

// ---------------------------------------------------------------------------
// device buffers and queries
// ---------------------------------------------------------------------------

DevBuf::~DevBuf()
{
    if (p) (void)hipFree(p);
}

int DevBuf::alloc(size_t n)
{
    if (p) { (void)hipFree(p); p = nullptr; }
    bytes = 0;
    if (!n) return OK;
    if (hipMalloc(&p, n) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return E_NOMEM;
    }
    bytes = n;
    return OK;
}

Plan::~Plan()
{
    for (auto e : prof_ev) (void)hipEventDestroy(e);
    r2r_destroy(r2r);
    rfp_destroy(rfp);
}

size_t Plan::device_bytes() const
{
    size_t b = s0.bytes + s1.bytes + tmp.bytes;
    for (auto &t : tables) b += t->bytes;
    return b;
}

int max_lds_bytes()
{
    static int cached = 0;
    if (!cached) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            cached = v;
        else {
            (void)hipGetLastError();
            cached = 64 * 1024;
        }
    }
    return cached;
}

int device_check(Error &err)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        err = {E_NODEVICE, "no HIP device available"};
        return E_NODEVICE;
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        err = {E_NODEVICE, "cannot query the current HIP device"};
        return E_NODEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        err = {E_NODEVICE, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950"};
        return E_NODEVICE;
    }
    return OK;
}

// ---------------------------------------------------------------------------
// arithmetic helpers
// ---------------------------------------------------------------------------

static long long env_ll(const char *name, long long def)
{
    const char *v = std::getenv(name);
    if (!v || !*v) return def;
    char *end = nullptr;
    const long long x = std::strtoll(v, &end, 10);
    return (end && *end == 0 && x > 0) ? x : def;
}

static int ilog2(long long v)
{
    int l = 0;
    while ((1LL << l) < v) ++l;
    return l;
}

static bool smooth13(long long n)
{
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

static std::vector<int> radices_for(long long n)
{
    std::vector<int> r;
    int twos = 0;
    while (n % 2 == 0) { n /= 2; ++twos; }
    while (twos >= 4) { r.push_back(16); twos -= 4; }
    if (twos == 3) r.push_back(8);
    else if (twos == 2) r.push_back(4);
    else if (twos == 1) r.push_back(2);
    for (int p : {3, 5, 7, 11, 13})
        while (n % p == 0) { r.push_back(p); n /= p; }
    return r;
}

static long long good_size7(long long n)
{
    long long best = LLONG_MAX;
    for (long long f7 = 1; f7 < best; f7 *= 7)
        for (long long f5 = f7; f5 < best; f5 *= 5)
            for (long long f3 = f5; f3 < best; f3 *= 3) {
                long long v = f3;
                while (v < n) v *= 2;
                if (v < best) best = v;
                if (f3 >= n) break;
            }
    return best;
}

static long long bluestein_n2(long long L)
{
    const long long need = 2 * L - 1;
    const long long s = good_size7(need);
    long long p2 = 1;
    while (p2 < need) p2 *= 2;
    return (p2 * 100 <= s * 115) ? p2 : s;
}

static int threads_per_transform(long long N, const std::vector<int> &rad);

// a length one LDS pass can take: 13-smooth, fits LDS and the workgroup size cap
static bool lds_len_ok(long long n)
{
    return n <= kMaxLdsLen && smooth13(n) && threads_per_transform(n, radices_for(n)) <= kMaxThreads;
}

// L = L1 * L2 with both factors single-pass lengths; L2 <= L1, as balanced as possible
static bool split_four(long long L, long long &L1, long long &L2)
{
    L1 = L2 = 0;
    for (long long a = 2; a * a <= L; ++a) {
        if (L % a) continue;
        const long long b = L / a;
        if (lds_len_ok(a) && lds_len_ok(b)) { L2 = a; L1 = b; }
    }
    return L2 > 0;
}

static int threads_per_transform(long long N, const std::vector<int> &rad)
{
    int nt = 1;
    for (int r : rad) {
        const long long nb = N / r;
        const int per = nb_max(r);
        nt = std::max<int>(nt, (int)((nb + per - 1) / per));
    }
    return nt;
}

// ---------------------------------------------------------------------------
// jobs
// ---------------------------------------------------------------------------

enum JobKind { J_C2C = 0, J_R2C = 1, J_C2R = 2 };

struct Job {
    int kind;
    long long L;         // transform length (real length for r2c / c2r)
    int forward;
    int adj = 0;         // adjoint weights on the real pre/post-processing
    bool fct;
    BufRef src, dst;
    long long is_ax, os_ax;                       // scalar-unit strides along the axis
    std::vector<long long> cnt, istr, ostr;       // batch dims, fastest first
    long long nbatch() const
    {
        long long b = 1;
        for (auto c : cnt) b *= c;
        return b;
    }
};

struct Builder {
    Plan &plan;
    Error &err;

    std::shared_ptr<DevBuf> upload(const std::vector<double2> &v)
    {
        auto b = std::make_shared<DevBuf>();
        if (b->alloc(v.size() * sizeof(double2)) != OK) {
            err = {E_NOMEM, "device allocation for a twiddle table failed"};
            return nullptr;
        }
        if (hipMemcpy(b->p, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "hipMemcpy of a twiddle table failed"};
            return nullptr;
        }
        plan.tables.push_back(b);
        return b;
    }

    // float32 plans keep the user's arrays in float / complex float; the
    // workspaces and tables stay FP64
    bool user32(BufRef r) const { return plan.dtype == 1 && (r == REF_IN || r == REF_OUT); }
    void io_flags(FftParams &P, BufRef src, BufRef dst) const
    {
        P.src32 = user32(src) ? 1 : 0;
        P.dst32 = user32(dst) ? 1 : 0;
    }

    const double2 *table_direct(long long n)   // W_n^j, j < n
    {
        std::vector<double2> v;
        twiddle_table(v, (uint64_t)n, (uint64_t)n, 1);
        auto b = upload(v);
        return b ? (const double2 *)b->p : nullptr;
    }

    const double2 *table_real(long long L)   // W_L^k, k <= L/2
    {
        std::vector<double2> v;
        twiddle_table(v, (uint64_t)L, (uint64_t)(L / 2 + 1), 1);
        auto b = upload(v);
        return b ? (const double2 *)b->p : nullptr;
    }

    bool table_two_level(long long L, const double2 *&lo, const double2 *&hi, long long &S)
    {
        S = 1;
        while (S * S < L) S *= 2;   // a power of two: the kernels split m with a shift and a mask
        std::vector<double2> a, b;
        twiddle_table(a, (uint64_t)L, (uint64_t)S, 1);
        twiddle_table(b, (uint64_t)L, (uint64_t)((L + S - 1) / S + 1), (uint64_t)S);
        auto ba = upload(a);
        if (!ba) return false;
        auto bb = upload(b);
        if (!bb) return false;
        lo = (const double2 *)ba->p;
        hi = (const double2 *)bb->p;
        return true;
    }

    static FftParams base(long long N)
    {
        FftParams P;
        std::memset(&P, 0, sizeof(P));
        P.N = (int)N;
        auto rad = radices_for(N);
        P.nst = (int)rad.size();
        for (size_t s = 0; s < rad.size(); ++s) P.radix[s] = rad[s];
        P.NT = threads_per_transform(N, rad);
        P.LD = (int)N;
        P.ilim = P.olim = LLONG_MAX;
        P.nload = P.nstore = (int)N;
        P.fct = 1.0;
        P.tw_mul = 1;
        P.sign = -1;
        P.ni = 1;
        return P;
    }

    static void side_dims(Side &s, const Job &J, bool input, bool pair)
    {
        std::memset(&s, 0, sizeof(s));
        s.kind = pair ? SIDE_DIMS_PAIR : SIDE_DIMS;
        s.nbd = (int)J.cnt.size();
        s.nphys = 1;
        for (int d = 0; d < s.nbd; ++d) {
            s.cnt[d] = J.cnt[d];
            s.str[d] = input ? J.istr[d] : J.ostr[d];
            s.nphys *= J.cnt[d];
        }
        s.pst = 1;
        s.sax = input ? J.is_ax : J.os_ax;
    }

    static void side_lin(Side &s, long long lin, long long ipos, long long pst)
    {
        std::memset(&s, 0, sizeof(s));
        s.kind = SIDE_LINEAR;
        s.lin = lin;
        s.ipos = ipos;
        s.pst = pst;
        s.sax = 1;
    }

    int choose_tb(const FftParams &P, long long ntf, bool col)
    {
        const long long per = (long long)P.LD * (long long)sizeof(double2);
        long long budget = col ? 128 * 1024 : 64 * 1024;
        const long long cap = max_lds_bytes() - (long long)sizeof(long long) * 8 * kMaxTB - 1024;
        budget = std::min(budget, cap);
        long long tb = std::max<long long>(1, budget / per);
        tb = std::min<long long>(tb, std::max(1, kMaxThreads / P.NT));
        tb = std::min<long long>(tb, kMaxTB);
        const long long want = std::max<long long>(1, (ntf + 1023) / 1024);
        if (tb > want) tb = std::max<long long>(want, col ? std::min<long long>(tb, 4) : 1);
        return (int)tb;
    }

    // the compile-time power-of-two kernel takes per-point load/store modes
    bool try_pow2(FftParams P, BufRef src, BufRef dst, bool uses_fct, bool colL, bool colS)
    {
        // an even-length r2c signal per transform (real post-process fused into the store) runs on
        // the row mapping; its N + 1 outputs are not a transform's N points
        const bool rhalf = P.store_mode == ST_R2C_HALF && P.load_mode == LD_REAL_HALF && P.nstore == P.N + 1;
        if (!pow2_supported(P.N) || P.conv || P.nload != P.N || (P.nstore != P.N && !rhalf)) return false;
        if ((P.store_mode == ST_R2C_HALF && !rhalf) || P.store_mode == ST_R2C_PAIR) return false;
        int map;
        if (colL && colS) map = 1;
        else if (!colL && colS) map = 2;
        else if (!colL && !colS) map = 0;
        else return false;
        if (rhalf && map != 0) return false;
        const int NT = P.N / 16;
        const long long per = (long long)pow2_lds_bytes(P.N, map);
        const long long cap = max_lds_bytes() - (long long)sizeof(long long) * 8 * kMaxTB - 1024;
        long long tb;
        // the kernel fixes its transforms per workgroup: 256 threads along rows,
        // 512 across transforms
        tb = pow2_tb_host(P.N, map);
        if (tb * per > cap) return false;
        P.TB = (int)tb;
        P.NT = NT;
        Pass ps;
        ps.kind = Pass::FFT_POW2;
        ps.map = map;
        ps.fp = P;
        ps.block = (unsigned)(tb * NT);
        const long long grid = (P.ntf + tb - 1) / tb;
        if (grid > INT_MAX) return false;
        ps.grid = (unsigned)std::max<long long>(1, grid);
        ps.lds = (size_t)(per * tb);
        ps.src = src;
        ps.dst = dst;
        ps.uses_fct = uses_fct;
        ps.name = "k_fft_pow2";
        ps.bytes_in = side_bytes(P, true);
        ps.bytes_out = side_bytes(P, false);
        make_alt(ps);
        plan.passes.push_back(ps);
        return true;
    }

    // LD_REAL_HALF with unit position stride reads x[2 pos], x[2 pos + 1]: the
    // same bytes as one complex word, so the pass can run as a plain complex
    // load when all batch strides are even and the pointer is 16-byte aligned
    static void make_alt(Pass &ps)
    {
        const FftParams &P = ps.fp;
        if (P.load_mode != LD_REAL_HALF || P.in.sax != 1 || P.in.kind != SIDE_DIMS || P.in.isig % 2) return;
        for (int d = 0; d < P.in.nbd; ++d)
            if (P.in.str[d] % 2) return;
        ps.fp_alt = P;
        ps.fp_alt.load_mode = LD_C2C;
        for (int d = 0; d < P.in.nbd; ++d) ps.fp_alt.in.str[d] /= 2;
        ps.fp_alt.in.isig /= 2;
        ps.has_alt = true;
    }

    // colL / colS: whether consecutive transforms are closer in memory than
    // consecutive points on the load / store side
    bool add_fft(FftParams P, BufRef src, BufRef dst, bool uses_fct, bool colL, bool colS)
    {
        io_flags(P, src, dst);
        if (try_pow2(P, src, dst, uses_fct, colL, colS)) return true;
        const long long per = (long long)P.LD * (long long)sizeof(double2) * P.TB;
        const long long cap = max_lds_bytes() - (long long)sizeof(long long) * 8 * kMaxTB - 1024;
        if (per > cap) {
            err = {E_UNSUPPORTED, "transform does not fit in LDS"};
            return false;
        }
        Pass ps;
        ps.kind = Pass::FFT_LDS;
        ps.fp = P;
        ps.block = (unsigned)(P.TB * P.NT);
        const long long grid = (P.ntf + P.TB - 1) / P.TB;
        if (grid > INT_MAX) {
            err = {E_UNSUPPORTED, "batch too large for one launch"};
            return false;
        }
        ps.grid = (unsigned)std::max<long long>(1, grid);
        ps.lds = (size_t)per;
        ps.src = src;
        ps.dst = dst;
        ps.uses_fct = uses_fct;
        ps.name = "k_fft_lds";
        ps.bytes_in = side_bytes(P, true);
        ps.bytes_out = side_bytes(P, false);
        plan.passes.push_back(ps);
        return true;
    }

    // bytes one launch reads (input) or writes (output) if each byte moves once
    // (float32 arrays move half the bytes of the FP64 figures below)
    static size_t side_bytes(const FftParams &P, bool input)
    {
        const size_t b = side_bytes64(P, input);
        return (input ? P.src32 : P.dst32) ? b / 2 : b;
    }
    static size_t side_bytes64(const FftParams &P, bool input)
    {
        long long n;   // points per logical transform on that side
        if (input) {
            const long long pts = std::min<long long>(P.nload, P.ilim);
            switch (P.load_mode) {
                case LD_REAL_HALF: return (size_t)(P.ntf * pts * 16);
                case LD_R2C_PAIR: return (size_t)(P.ntf * pts * 16);
                case LD_C2R_HALF: n = (P.ni > 1) ? pts : (long long)P.M + 1; return (size_t)(P.ntf * n * 16);
                case LD_C2R_PAIR: n = (P.ni > 1) ? pts : (long long)P.M / 2 + 1; return (size_t)(P.ntf * n * 32);
                default: return (size_t)(P.ntf * pts * 16);
            }
        }
        const long long pts = std::min<long long>(P.nstore, P.olim);
        switch (P.store_mode) {
            case ST_R2C_PAIR: return (size_t)(P.ntf * pts * 32);
            case ST_C2R_HALF: return (size_t)(P.ntf * pts * 16);
            case ST_C2R_PAIR: return (size_t)(P.ntf * pts * 16);
            default: return (size_t)(P.ntf * pts * 16);
        }
    }

    // ---- one LDS pass over the whole job --------------------------------
    bool single_pass(const Job &J)
    {
        const long long L = J.L;
        const bool odd = (L & 1) != 0;
        const bool real = J.kind != J_C2C;
        const bool pair = real && odd;
        const long long M = real ? (odd ? L : L / 2) : L;    // complex points in LDS
        FftParams P = base(M);
        P.tw = M > 1 ? table_direct(M) : nullptr;
        if (M > 1 && !P.tw) return false;
        P.fwd_flag = J.forward;
        P.adj_w = J.adj;
        P.M = (int)M;
        side_dims(P.in, J, true, pair);
        side_dims(P.out, J, false, pair);
        P.ntf = pair ? (J.nbatch() + 1) / 2 : J.nbatch();
        switch (J.kind) {
            case J_C2C:
                P.load_mode = LD_C2C; P.store_mode = ST_C2C;
                P.sign = J.forward ? -1 : 1;
                P.olim = L;
                break;
            case J_R2C:
                P.sign = -1;
                if (pair) { P.load_mode = LD_R2C_PAIR; P.store_mode = ST_R2C_PAIR; P.nstore = (int)(L / 2 + 1); }
                else {
                    P.load_mode = LD_REAL_HALF; P.store_mode = ST_R2C_HALF; P.nstore = (int)(M + 1);
                    P.rtw = table_real(L);
                    if (!P.rtw) return false;
                }
                break;
            default:
                P.sign = 1;
                P.olim = M;
                if (pair) { P.load_mode = LD_C2R_PAIR; P.store_mode = ST_C2R_PAIR; }
                else {
                    P.load_mode = LD_C2R_HALF; P.store_mode = ST_C2R_HALF;
                    P.rtw = table_real(L);
                    if (!P.rtw) return false;
                }
                break;
        }
        const bool colL = !J.cnt.empty() && std::llabs(J.istr[0]) < std::llabs(J.is_ax);
        const bool colS = !J.cnt.empty() && std::llabs(J.ostr[0]) < std::llabs(J.os_ax);
        P.TB = choose_tb(P, P.ntf, colL || colS);
        P.col_load = P.TB > 1 && colL;
        P.col_store = P.TB > 1 && colS;
        return add_fft(P, J.src, J.dst, J.fct, colL, colS);
    }

    // ---- Bluestein tables ------------------------------------------------
    bool bluestein_tables(long long L, long long n2, const double2 *&chirp, const double2 *&spec)
    {
        // w[n] = exp(-i pi n^2 / L) = W_{2L}^{n^2 mod 2L}
        std::vector<double2> w(L), b(n2, make_double2(0.0, 0.0));
        for (long long n = 0; n < L; ++n) {
            const uint64_t idx = (uint64_t)(((unsigned __int128)n * (unsigned __int128)n) % (unsigned __int128)(2 * L));
            w[n] = twiddle_at(idx, (uint64_t)(2 * L));
        }
        for (long long m = 0; m < L; ++m) {
            const double2 c = make_double2(w[m].x, -w[m].y);
            b[m] = c;
            if (m) b[n2 - m] = c;
        }
        auto bw = upload(w);
        if (!bw) return false;
        auto bs = upload(b);
        if (!bs) return false;
        chirp = (const double2 *)bw->p;
        // spectrum of b, scaled by 1/n2, computed in place on the GPU
        std::unique_ptr<Plan> sub;
        const size_t shp[1] = {(size_t)n2};
        const ptrdiff_t st[1] = {(ptrdiff_t)sizeof(double2)};
        const size_t ax[1] = {0};
        Error e2;
        if (plan_create(sub, J_C2C, 0, 1, shp, st, st, 1, ax, 1, e2) != OK) { err = e2; return false; }
        if (plan_execute(*sub, bs->p, bs->p, 1.0 / (double)n2, nullptr, e2) != OK) { err = e2; return false; }
        if (hipDeviceSynchronize() != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "Bluestein spectrum computation failed"};
            return false;
        }
        spec = (const double2 *)bs->p;
        return true;
    }

    bool bluestein_single(const Job &J, long long Ls, long long n2)
    {
        // Ls: complex signal length the chirp-z runs on (L, L/2 for even real)
        const double2 *chirp = nullptr, *spec = nullptr;
        if (!bluestein_tables(Ls, n2, chirp, spec)) return false;
        const bool real = J.kind != J_C2C;
        const bool pair = real && (J.L & 1);
        FftParams P = base(n2);
        P.tw = table_direct(n2);
        if (!P.tw) return false;
        P.chirp = chirp;
        P.spec = spec;
        P.chirp_in = P.conv = P.chirp_out = 1;
        P.sign = -1;
        P.fwd_flag = J.forward;
        P.adj_w = J.adj;
        P.M = (int)Ls;
        P.ilim = Ls;
        P.nload = (int)Ls;
        P.olim = Ls;
        side_dims(P.in, J, true, pair);
        side_dims(P.out, J, false, pair);
        P.ntf = pair ? (J.nbatch() + 1) / 2 : J.nbatch();
        switch (J.kind) {
            case J_C2C:
                P.load_mode = LD_C2C; P.store_mode = ST_C2C;
                P.conj_in = P.conj_out = J.forward ? 0 : 1;
                P.nstore = (int)Ls;
                break;
            case J_R2C:
                if (pair) { P.load_mode = LD_R2C_PAIR; P.store_mode = ST_R2C_PAIR; P.nstore = (int)(Ls / 2 + 1); }
                else {
                    P.load_mode = LD_REAL_HALF; P.store_mode = ST_R2C_HALF; P.nstore = (int)(Ls + 1);
                    P.rtw = table_real(J.L);
                    if (!P.rtw) return false;
                }
                break;
            default:
                P.conj_in = P.conj_out = 1;
                P.nstore = (int)Ls;
                if (pair) { P.load_mode = LD_C2R_PAIR; P.store_mode = ST_C2R_PAIR; }
                else {
                    P.load_mode = LD_C2R_HALF; P.store_mode = ST_C2R_HALF;
                    P.rtw = table_real(J.L);
                    if (!P.rtw) return false;
                }
                break;
        }
        const bool colL = !J.cnt.empty() && std::llabs(J.istr[0]) < std::llabs(J.is_ax);
        const bool colS = !J.cnt.empty() && std::llabs(J.ostr[0]) < std::llabs(J.os_ax);
        P.TB = choose_tb(P, P.ntf, colL || colS);
        P.col_load = P.TB > 1 && colL;
        P.col_store = P.TB > 1 && colS;
        return add_fft(P, J.src, J.dst, J.fct, colL, colS);
    }

    // ---- four-step (and four-step Bluestein) -----------------------------
    bool ensure_ws(size_t &have, size_t need)
    {
        have = std::max(have, need);
        return true;
    }

    bool four_step(const Job &J)
    {
        const bool real = J.kind != J_C2C;
        const bool pair = real && (J.L & 1);
        const long long Ls = real ? ((J.L & 1) ? J.L : J.L / 2) : J.L;   // complex signal length
        const bool blue = !smooth13(Ls);
        const long long n2 = blue ? bluestein_n2(Ls) : Ls;
        long long L1, L2;
        if (!split_four(n2, L1, L2)) {
            err = {E_UNSUPPORTED, "transform length too long for two passes"};
            return false;
        }
        const long long nlog = pair ? (J.nbatch() + 1) / 2 : J.nbatch();
        // twiddles: W_{n2} two-level for the pass-A twiddle, W_{2 Ls} two-level
        // for the even-real pre/post twiddle
        const bool real_half = real && !pair;
        const double2 *lo, *hi;
        long long S;
        if (!table_two_level(n2, lo, hi, S)) return false;
        const double2 *rlo = nullptr, *rhi = nullptr;
        long long rS = 0;
        if (real_half && !table_two_level(2 * Ls, rlo, rhi, rS)) return false;
        const double2 *twA = table_direct(L2), *twB = table_direct(L1);
        if (!twA || !twB) return false;
        const double2 *chirp = nullptr, *spec = nullptr, *spec_rows = nullptr;
        if (blue && !bluestein_tables(Ls, n2, chirp, spec)) return false;

        // Optional batch chunks (IMPULSE_FFT_MALL_MB): the workspace of one chunk
        // stays within the budget so that pass A's output could still sit in the
        // MALL (Infinity Cache) when pass B reads it.  Measured on MI355X the
        // smaller launches lose more than the cache saves, so it is off unless set.
        const long long mall = env_ll("IMPULSE_FFT_MALL_MB", 1LL << 30) * (1LL << 20);
        const long long per_sig = n2 * (long long)sizeof(double2);
        const long long chunk = std::max<long long>(1, std::min<long long>(nlog, mall / per_sig));
        long long cn = chunk;   // signals in the current chunk (captured by the pass builders)
        const size_t ws = (size_t)(chunk * n2) * sizeof(double2);
        ensure_ws(plan.s0_bytes, ws);
        ensure_ws(plan.s1_bytes, ws);

        auto passA = [&](FftParams P) {
            P.N = (int)L2;
            auto rad = radices_for(L2);
            P.nst = (int)rad.size();
            for (size_t s = 0; s < rad.size(); ++s) P.radix[s] = rad[s];
            P.NT = threads_per_transform(L2, rad);
            P.LD = (int)L2;
            P.tw = twA;
            P.ni = L1;
            P.ntf = L1 * cn;
            P.nload = P.nstore = (int)L2;
            P.store_mode = ST_TWIDDLE;
            P.tw_lo = lo; P.tw_hi = hi; P.tw_S = S; P.tw_L = n2; P.tw_mul = 1; P.tw_logS = ilog2(S);
            P.rtw_lo = rlo; P.rtw_hi = rhi; P.rtw_S = rS; P.rtw_logS = rS ? ilog2(rS) : 0;
            side_lin(P.out, n2, 1, L1);
            P.TB = choose_tb(P, P.ntf, true);
            P.col_load = P.col_store = P.TB > 1;
            return P;
        };
        auto passB = [&](FftParams P) {
            P.N = (int)L1;
            auto rad = radices_for(L1);
            P.nst = (int)rad.size();
            for (size_t s = 0; s < rad.size(); ++s) P.radix[s] = rad[s];
            P.NT = threads_per_transform(L1, rad);
            P.LD = (int)L1;
            P.tw = twB;
            P.ni = L2;
            P.ntf = L2 * cn;
            P.nload = P.nstore = (int)L1;
            P.load_mode = LD_C2C;
            side_lin(P.in, n2, L1, 1);
            P.TB = choose_tb(P, P.ntf, true);
            P.col_load = 0;
            P.col_store = P.TB > 1;
            return P;
        };

        FftParams P0 = base(1);
        P0.fwd_flag = J.forward;
        P0.adj_w = J.adj;
        P0.M = (int)Ls;
        // input side of pass A: the user's layout, inner index = n1 (position step L1)
        side_dims(P0.in, J, true, pair);
        P0.in.ipos = 1;
        P0.in.pst = L1;
        // output side of the final pass
        Side outS;
        side_dims(outS, J, false, pair);
        outS.ipos = 1;
        outS.pst = L2;

        int lmode;
        int dir_sign;   // sign of the FFT proper
        switch (J.kind) {
            case J_C2C: lmode = LD_C2C; dir_sign = J.forward ? -1 : 1; break;
            case J_R2C: lmode = pair ? LD_R2C_PAIR : LD_REAL_HALF; dir_sign = -1; break;
            default: lmode = pair ? LD_C2R_PAIR : LD_C2R_HALF; dir_sign = 1; break;
        }
        // real-half load twiddle W_{2 Ls}^{-p} is read from the two-level table
        P0.rtw = nullptr;

        const int conj = (J.kind == J_C2R) || (J.kind == J_C2C && !J.forward);
        auto emit = [&](long long u0) -> bool {
            P0.in.u0 = u0;
            outS.u0 = u0;
            if (!blue) {
                FftParams A = passA(P0);
                A.load_mode = lmode;
                A.sign = dir_sign;
                if (!add_fft(A, J.src, REF_S0, false, true, true)) return false;
                FftParams B = passB(P0);
                B.sign = dir_sign;
                if (J.kind == J_R2C) {
                    // even lengths: the real post-process runs inside pass B on row pairs
                    if (!pair && rowpair_ok(L1, L2, ws)) {
                        B.store_mode = ST_R2C_ROWPAIR;
                        side_dims(B.out, J, false, false);
                        B.out.u0 = u0;
                        B.M = (int)Ls;
                        B.rtw_lo = rlo; B.rtw_hi = rhi; B.rtw_S = rS; B.rtw_logS = ilog2(rS);
                        return add_rowpair(B, REF_S0, J.dst, J.fct, cn, Ls);
                    }
                    side_lin(B.out, n2, 1, L2);
                    B.store_mode = ST_C2C;
                    if (!add_fft(B, REF_S0, REF_S1, false, false, true)) return false;
                    return add_post(J, pair, Ls, n2, rlo, rhi, rS, u0, cn);
                }
                B.out = outS;
                B.store_mode = (J.kind == J_C2C) ? ST_C2C : (pair ? ST_C2R_PAIR : ST_C2R_HALF);
                B.olim = Ls;
                return add_fft(B, REF_S0, J.dst, J.fct, false, true);
            }

            if (conv_rows_ok(L1, n2, ws)) {
                // Bluestein in three passes: A1 (chirp in, forward columns, twiddle) ->
                // B (each row in place: forward, times the spectrum, inverse, conjugate
                // twiddle) -> A2 (inverse columns, chirp out, natural order).  With
                // n = j + L1 n' and the spectrum index f = k + L2 k', the inverse
                // transform splits the same way as the forward one, so the spectrum
                // never has to be stored in natural order.
                if (!spec_rows && !(spec_rows = spectrum_rows(spec, L1, L2))) return false;
                FftParams A1 = passA(P0);
                A1.load_mode = lmode;
                A1.sign = -1;
                A1.chirp = chirp;
                A1.chirp_in = 1;
                A1.conj_in = conj;
                A1.ilim = Ls;
                if (!add_fft(A1, J.src, REF_S0, false, true, true)) return false;
                FftParams B = passB(P0);
                B.sign = -1;
                B.store_mode = ST_CONV_TWIDDLE;
                B.mtab = spec_rows;
                side_lin(B.out, n2, L1, 1);
                B.tw_lo = lo; B.tw_hi = hi; B.tw_S = S; B.tw_L = n2; B.tw_mul = 1; B.tw_logS = ilog2(S);
                if (!add_conv_rows(B, REF_S0, REF_S0)) return false;
                FftParams P1 = P0;
                side_lin(P1.in, n2, 1, L1);
                FftParams A2 = passA(P1);
                A2.load_mode = LD_C2C;
                A2.sign = 1;
                A2.chirp = chirp;
                A2.chirp_out = 1;
                A2.conj_out = conj;
                A2.olim = Ls;
                if (J.kind == J_R2C) {
                    A2.store_mode = ST_C2C;   // natural order into the workspace, then the real post-process
                    if (!add_fft(A2, REF_S0, REF_S1, false, true, true)) return false;
                    return add_post(J, pair, Ls, n2, rlo, rhi, rS, u0, cn);
                }
                Side outA;
                side_dims(outA, J, false, pair);
                outA.ipos = 1;
                outA.pst = L1;
                outA.u0 = u0;
                A2.out = outA;
                A2.store_mode = (J.kind == J_C2C) ? ST_C2C : (pair ? ST_C2R_PAIR : ST_C2R_HALF);
                return add_fft(A2, REF_S0, J.dst, J.fct, true, true);
            }

            // Bluestein: A1 (chirp in) -> B1 (x spectrum) -> A2 (inverse) -> B2 (chirp out)
            FftParams A1 = passA(P0);
            A1.load_mode = lmode;
            A1.sign = -1;
            A1.chirp = chirp;
            A1.chirp_in = 1;
            A1.conj_in = conj;
            A1.ilim = Ls;
            if (!add_fft(A1, J.src, REF_S0, false, true, true)) return false;
            FftParams B1 = passB(P0);
            B1.sign = -1;
            B1.store_mode = ST_MUL_TABLE;
            B1.mtab = spec;
            side_lin(B1.out, n2, 1, L2);
            if (!add_fft(B1, REF_S0, REF_S1, false, false, true)) return false;
            FftParams P1 = P0;
            side_lin(P1.in, n2, 1, L1);
            FftParams A2 = passA(P1);
            A2.load_mode = LD_C2C;
            A2.sign = 1;
            if (!add_fft(A2, REF_S1, REF_S0, false, true, true)) return false;
            FftParams B2 = passB(P0);
            B2.sign = 1;
            B2.chirp = chirp;
            B2.chirp_out = 1;
            B2.conj_out = conj;
            B2.olim = Ls;
            if (J.kind == J_R2C) {
                side_lin(B2.out, n2, 1, L2);
                B2.store_mode = ST_C2C;
                if (!add_fft(B2, REF_S0, REF_S1, false, false, true)) return false;
                return add_post(J, pair, Ls, n2, rlo, rhi, rS, u0, cn);
            }
            B2.out = outS;
            B2.store_mode = (J.kind == J_C2C) ? ST_C2C : (pair ? ST_C2R_PAIR : ST_C2R_HALF);
            return add_fft(B2, REF_S0, J.dst, J.fct, false, true);
        };
        for (long long u0 = 0; u0 < nlog; u0 += chunk) {
            cn = std::min(chunk, nlog - u0);
            if (!emit(u0)) return false;
        }
        return true;
    }

    // the three-pass Bluestein: the row pass runs on the power-of-two row kernel
    // (rows of L1 points), whose buffer loads address the workspace and the
    // spectrum table with 31-bit byte offsets
    static bool conv_rows_ok(long long L1, long long n2, size_t ws)
    {
        if (env_ll("IMPULSE_FFT_NO_CONV_ROWS", 0)) return false;
        return pow2_supported((int)L1) && L1 >= 64 && ws <= (size_t)0x7fffffff && n2 * 16 <= 0x7fffffffLL;
    }

    // the kernel spectrum in row order of the four-step split n2 = L1 L2:
    // entry k L1 + k' holds spectrum bin k + L2 k' (k < L2, k' < L1)
    const double2 *spectrum_rows(const double2 *spec, long long L1, long long L2)
    {
        const long long n2 = L1 * L2;
        std::vector<double2> nat((size_t)n2), rows((size_t)n2);
        if (hipMemcpy(nat.data(), spec, (size_t)n2 * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) {
            (void)hipGetLastError();
            err = {E_HIP, "hipMemcpy of the Bluestein spectrum failed"};
            return nullptr;
        }
        for (long long k = 0; k < L2; ++k)
            for (long long kp = 0; kp < L1; ++kp) rows[(size_t)(k * L1 + kp)] = nat[(size_t)(k + L2 * kp)];
        auto b = upload(rows);
        return b ? (const double2 *)b->p : nullptr;
    }

    bool add_conv_rows(FftParams P, BufRef src, BufRef dst)
    {
        io_flags(P, src, dst);
        const int map = 0;
        const long long tb = pow2_tb_host(P.N, map);
        const long long per = (long long)pow2_lds_bytes(P.N, map);
        const long long cap = max_lds_bytes() - (long long)sizeof(long long) * 8 * kMaxTB - 1024;
        if (tb * per > cap) {
            err = {E_UNSUPPORTED, "transform does not fit in LDS"};
            return false;
        }
        P.TB = (int)tb;
        P.NT = P.N / 16;
        P.col_load = P.col_store = 0;
        Pass ps;
        ps.kind = Pass::FFT_POW2;
        ps.map = map;
        ps.fp = P;
        ps.block = (unsigned)(tb * P.NT);
        const long long grid = (P.ntf + tb - 1) / tb;
        if (grid > INT_MAX) {
            err = {E_UNSUPPORTED, "batch too large for one launch"};
            return false;
        }
        ps.grid = (unsigned)std::max<long long>(1, grid);
        ps.lds = (size_t)(per * tb);
        ps.src = src;
        ps.dst = dst;
        ps.uses_fct = false;
        ps.name = "k_fft_pow2";
        ps.bytes_in = side_bytes(P, true) + (size_t)(P.ntf * P.N * 16);   // rows and the spectrum table
        ps.bytes_out = side_bytes(P, false);
        plan.passes.push_back(ps);
        return true;
    }

    // pass B with the fused real post-process: the power-of-two row kernel with
    // two or four row pairs per workgroup, loads from a workspace addressable
    // with 31-bit byte offsets
    static bool rowpair_ok(long long L1, long long L2, size_t ws)
    {
        if (env_ll("IMPULSE_FFT_NO_ROWPAIR", 0)) return false;
        const int tb = (L1 == 1024 || L1 == 2048) ? pow2_tb_host((int)L1, 2) : 0;
        return (tb == 4 || tb == 8) && L2 >= 4 && L2 % 2 == 0 && ws <= (size_t)0x7fffffff;
    }

    bool add_rowpair(FftParams P, BufRef src, BufRef dst, bool uses_fct, long long signals, long long Ls)
    {
        io_flags(P, src, dst);
        const int map = 2;
        const long long tb = pow2_tb_host(P.N, map);
        const long long per = (long long)pow2_lds_bytes(P.N, map);
        const long long cap = max_lds_bytes() - (long long)sizeof(long long) * 8 * kMaxTB - 1024;
        if (tb * per > cap) {
            err = {E_UNSUPPORTED, "transform does not fit in LDS"};
            return false;
        }
        P.TB = (int)tb;
        P.NT = P.N / 16;
        Pass ps;
        ps.kind = Pass::FFT_POW2;
        ps.map = map;
        ps.fp = P;
        ps.block = (unsigned)(tb * P.NT);
        const long long grid = pow2_rowpair_grid(signals, P.ni, (int)tb);
        if (grid > INT_MAX) {
            err = {E_UNSUPPORTED, "batch too large for one launch"};
            return false;
        }
        ps.grid = (unsigned)grid;
        ps.lds = (size_t)(per * tb);
        ps.src = src;
        ps.dst = dst;
        ps.uses_fct = uses_fct;
        ps.name = "k_fft_pow2";
        ps.bytes_in = (size_t)(P.ntf * P.N * 16);
        ps.bytes_out = (size_t)(signals * (Ls + 1) * (P.dst32 ? 8 : 16));
        plan.passes.push_back(ps);
        return true;
    }

    bool add_post(const Job &J, bool pair, long long Ls, long long zlen, const double2 *lo, const double2 *hi,
                  long long S, long long u0, long long cn)
    {
        Pass ps;
        ps.kind = Pass::REAL_POST;
        RealPostParams &R = ps.rp;
        std::memset(&R, 0, sizeof(R));
        R.half = pair ? 0 : 1;
        R.M = (int)Ls;
        R.fwd_flag = J.forward;
        R.adj_w = J.adj;
        R.dst32 = user32(J.dst) ? 1 : 0;
        R.fct = 1.0;
        R.ntf = cn;
        R.zlen = zlen;
        side_dims(R.out, J, false, pair);
        R.out.u0 = u0;
        R.two_level = 1;
        R.tw_lo = lo;
        R.tw_hi = hi;
        R.tw_S = S;
        R.tw_logS = ilog2(S);
        ps.src = REF_S1;
        ps.dst = J.dst;
        ps.uses_fct = J.fct;
        ps.name = "k_real_post";
        const long long per = pair ? Ls / 2 + 1 : Ls + 1;
        ps.bytes_in = (size_t)(R.ntf * Ls * 16);
        ps.bytes_out = (size_t)(R.ntf * per * (pair ? 32 : 16)) / (R.dst32 ? 2 : 1);
        plan.passes.push_back(ps);
        return true;
    }

    bool build_job(const Job &J)
    {
        const bool real = J.kind != J_C2C;
        const long long Ls = real ? ((J.L & 1) ? J.L : J.L / 2) : J.L;
        if (lds_len_ok(Ls)) return single_pass(J);
        if (!smooth13(Ls)) {
            const long long n2 = bluestein_n2(Ls);
            if (lds_len_ok(n2)) return bluestein_single(J, Ls, n2);
        }
        return four_step(J);
    }
};

// ---------------------------------------------------------------------------
// plan creation
// ---------------------------------------------------------------------------

static bool merge_dims(Job &J)
{
    // drop unit dims, sort by |input stride|, merge dims that are contiguous in both arrays
    std::vector<size_t> idx;
    for (size_t d = 0; d < J.cnt.size(); ++d)
        if (J.cnt[d] > 1) idx.push_back(d);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](size_t a, size_t b) { return std::llabs(J.istr[a]) < std::llabs(J.istr[b]); });
    std::vector<long long> c, is, os;
    for (size_t d : idx) {
        if (!c.empty() && c.back() * is.back() == J.istr[d] && c.back() * os.back() == J.ostr[d]) {
            c.back() *= J.cnt[d];
            continue;
        }
        c.push_back(J.cnt[d]);
        is.push_back(J.istr[d]);
        os.push_back(J.ostr[d]);
    }
    J.cnt = c;
    J.istr = is;
    J.ostr = os;
    return J.cnt.size() <= (size_t)kMaxBatchDims;
}

static Job make_job(int kind, long long L, int forward, bool fct, BufRef src, BufRef dst, size_t axis,
                    const std::vector<size_t> &shape, const std::vector<long long> &is,
                    const std::vector<long long> &os)
{
    Job J;
    J.kind = kind;
    J.L = L;
    J.forward = forward;
    J.fct = fct;
    J.src = src;
    J.dst = dst;
    J.is_ax = is[axis];
    J.os_ax = os[axis];
    for (size_t d = 0; d < shape.size(); ++d) {
        if (d == axis) continue;
        J.cnt.push_back((long long)shape[d]);
        J.istr.push_back(is[d]);
        J.ostr.push_back(os[d]);
    }
    return J;
}

static void span(const std::vector<size_t> &shape, const std::vector<ptrdiff_t> &st, ptrdiff_t elem,
                 ptrdiff_t &lo, ptrdiff_t &hi)
{
    lo = 0;
    hi = 0;
    for (size_t d = 0; d < shape.size(); ++d) {
        const ptrdiff_t e = (ptrdiff_t)(shape[d] - 1) * st[d];
        if (e < 0) lo += e;
        else hi += e;
    }
    hi += elem;
}

// byte spans of a plan whose input and output have the plan's shape
void plan_spans(Plan &P, ptrdiff_t in_elem, ptrdiff_t out_elem)
{
    span(P.shape, P.sin, in_elem, P.in_lo, P.in_hi);
    span(P.shape, P.sout, out_elem, P.out_lo, P.out_hi);
    size_t n = 1;
    for (auto s : P.shape) n *= s;
    P.out_dense = (size_t)(P.out_hi - P.out_lo) == n * (size_t)out_elem;
}

int plan_create(std::unique_ptr<Plan> &out, int kind, int dtype, size_t ndim, const size_t *shape,
                const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                int forward, Error &err)
{
    if (kind < 0 || kind > 4) { err = {E_INVALID, "unknown transform kind"}; return E_INVALID; }
    // adjoints of the real transforms (with respect to the real inner product):
    // r2c* = c2r of the other direction with interior bins weighted 1/2,
    // c2r* = r2c of the other direction with interior bins weighted 2
    int adj = 0;
    if (kind == 3) { kind = J_C2R; forward = !forward; adj = 1; }
    else if (kind == 4) { kind = J_R2C; forward = !forward; adj = 1; }
    if (dtype != 0 && dtype != 1) { err = {E_INVALID, "unknown dtype"}; return E_INVALID; }
    if (ndim < 1 || !shape) { err = {E_INVALID, "ndim must be >= 1"}; return E_INVALID; }
    if (!stride_in || !stride_out) { err = {E_INVALID, "stride dimension mismatch"}; return E_INVALID; }
    if (naxes < 1 || !axes) { err = {E_INVALID, "no axes given"}; return E_INVALID; }
    std::vector<int> seen(ndim, 0);
    for (size_t a = 0; a < naxes; ++a) {
        if (axes[a] >= ndim) { err = {E_INVALID, "bad axis number"}; return E_INVALID; }
        if (++seen[axes[a]] > 1) { err = {E_INVALID, "axis specified repeatedly"}; return E_INVALID; }
    }
    int rc = device_check(err);
    if (rc != OK) return rc;

    auto p = std::make_unique<Plan>();
    Plan &P = *p;
    P.kind = kind;
    P.dtype = dtype;
    P.forward = forward ? 1 : 0;
    (void)hipGetDevice(&P.device);
    P.shape.assign(shape, shape + ndim);
    P.sin.assign(stride_in, stride_in + ndim);
    P.sout.assign(stride_out, stride_out + ndim);
    P.axes.assign(axes, axes + naxes);
    for (size_t d = 0; d < ndim; ++d)
        if (shape[d] == 0) P.empty = true;

    const ptrdiff_t es = dtype == 1 ? 4 : 8;   // bytes of one real scalar of the user's arrays
    const ptrdiff_t in_elem = (kind == J_R2C) ? es : 2 * es, out_elem = (kind == J_C2R) ? es : 2 * es;
    const size_t last = axes[naxes - 1];
    std::vector<size_t> shape_c(P.shape);   // complex-side shape for r2c / c2r
    if (kind != J_C2C) shape_c[last] = P.shape[last] / 2 + 1;
    const std::vector<size_t> &shape_in = (kind == J_C2R) ? shape_c : P.shape;
    const std::vector<size_t> &shape_out = (kind == J_R2C) ? shape_c : P.shape;
    if (P.empty) {
        out = std::move(p);
        return OK;
    }
    for (size_t d = 0; d < ndim; ++d) {
        if (P.sin[d] % in_elem || P.sout[d] % out_elem) {
            err = {E_UNSUPPORTED, "strides must be multiples of the element size"};
            return E_UNSUPPORTED;
        }
    }
    span(shape_in, P.sin, in_elem, P.in_lo, P.in_hi);
    span(shape_out, P.sout, out_elem, P.out_lo, P.out_hi);
    {
        size_t n = 1;
        for (auto s : shape_out) n *= s;
        P.out_dense = (size_t)(P.out_hi - P.out_lo) == n * (size_t)out_elem;
    }

    std::vector<long long> is(ndim), os(ndim);
    for (size_t d = 0; d < ndim; ++d) {
        is[d] = P.sin[d] / in_elem;
        os[d] = P.sout[d] / out_elem;
    }

    std::vector<Job> jobs;
    if (kind == J_C2C) {
        // The axes are separable, so their order changes only the rounding.  The axis with the
        // smallest input stride runs first: its pass reads the input and writes the output in
        // whole lines, and the strided passes that follow run in place on lines they have just
        // read, which stay in the caches for their stores.  Timed alone (20 executes of one plan),
        // c2c 4096^2 over [0, 1] took 0.297 ms columns first and 0.282 ms rows first, 2048 x 8192
        // 0.301 -> 0.285 ms and 8192^2 1.78 -> 1.71 ms; between other transforms the gain is
        // within run-to-run noise.
        std::vector<size_t> order(axes, axes + naxes);
        size_t lead = 0;
        for (size_t a = 1; a < naxes; ++a)
            if (std::llabs(is[order[a]]) < std::llabs(is[order[lead]])) lead = a;
        std::rotate(order.begin(), order.begin() + lead, order.begin() + lead + 1);
        for (size_t a = 0; a < naxes; ++a) {
            const bool first = a == 0;
            jobs.push_back(make_job(J_C2C, (long long)shape[order[a]], P.forward, first, first ? REF_IN : REF_OUT,
                                    REF_OUT, order[a], P.shape, first ? is : os, os));
        }
    } else if (kind == J_R2C) {
        jobs.push_back(make_job(J_R2C, (long long)shape[last], P.forward, true, REF_IN, REF_OUT, last, P.shape, is, os));
        jobs.back().adj = adj;
        for (size_t a = 0; a + 1 < naxes; ++a)
            jobs.push_back(make_job(J_C2C, (long long)shape_c[axes[a]], P.forward, false, REF_OUT, REF_OUT, axes[a],
                                    shape_c, os, os));
    } else {
        std::vector<long long> ts(ndim);   // contiguous complex temporary
        if (naxes > 1) {
            long long acc = 1;
            for (size_t d = ndim; d-- > 0;) {
                ts[d] = acc;
                acc *= (long long)shape_c[d];
            }
            P.tmp_bytes = (size_t)acc * 16;
            for (size_t a = 0; a + 1 < naxes; ++a) {
                const bool first = a == 0;
                jobs.push_back(make_job(J_C2C, (long long)shape_c[axes[a]], P.forward, false, first ? REF_IN : REF_TMP,
                                        REF_TMP, axes[a], shape_c, first ? is : ts, ts));
            }
        }
        const std::vector<long long> &cis = naxes > 1 ? ts : is;
        Job J = make_job(J_C2R, (long long)shape[last], P.forward, true, naxes > 1 ? REF_TMP : REF_IN, REF_OUT, last,
                         P.shape, cis, os);
        J.adj = adj;
        jobs.push_back(J);
    }

    Builder B{P, err};
    for (auto &J : jobs) {
        if (!merge_dims(J)) { err = {E_UNSUPPORTED, "too many batch dimensions"}; return E_UNSUPPORTED; }
        if (!B.build_job(J)) return err.code ? err.code : E_UNSUPPORTED;
    }
    if (P.s0.alloc(P.s0_bytes) != OK || P.s1.alloc(P.s1_bytes) != OK || P.tmp.alloc(P.tmp_bytes) != OK) {
        err = {E_NOMEM, "device workspace allocation failed"};
        return E_NOMEM;
    }
    out = std::move(p);
    return OK;
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------

// Workspace pointers of one execute: the plan's own buffers, or a borrowed region
// (impulse_fft_execute_noalloc).
struct Workspace {
    void *s0, *s1, *tmp;
    size_t b0, b1, bt;
};

static int plan_execute_ws(Plan &P, const void *in, void *out, double fct, hipStream_t stream, const Workspace &W,
                           Error &err);

int plan_execute(Plan &P, const void *in, void *out, double fct, hipStream_t stream, Error &err)
{
    const Workspace W{P.s0.p, P.s1.p, P.tmp.p, P.s0.bytes, P.s1.bytes, P.tmp.bytes};
    return plan_execute_ws(P, in, out, fct, stream, W, err);
}

// ---------------------------------------------------------------------------
// execution without allocation: the workspace comes from one static device
// array per GPU instead of the plan's buffers
// ---------------------------------------------------------------------------

// Every call on a GPU shares that GPU's copy of this array (it keeps the last
// call's contents; the passes write every workspace word before reading it), so
// calls that overlap on one GPU race.  The host uses the per-device address
// from hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mtx;
static void *g_scratch_ptr[kMaxDevices];

static int scratch_address(void **out, Error &err)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        err = {E_NODEVICE, "no usable current device for the no-allocation scratch"};
        return E_NODEVICE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            err = {E_HIP, std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                              hipGetErrorString(e)};
            return E_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int plan_execute_noalloc(Plan &P, const void *in, void *out, double fct, hipStream_t stream, Error &err)
{
    if (P.r2r || P.rfp) {
        err = {E_UNSUPPORTED, "no-allocation execution takes c2c / r2c / c2r plans"};
        return E_UNSUPPORTED;
    }
    const size_t need = align256(P.s0_bytes) + align256(P.s1_bytes) + align256(P.tmp_bytes);
    if (need > kNoAllocScratchBytes) {
        err = {E_UNSUPPORTED, "the plan's workspace (" + std::to_string(need) + " bytes) exceeds the " +
                                  std::to_string(kNoAllocScratchBytes) + "-byte no-allocation scratch"};
        return E_UNSUPPORTED;
    }
    char *base = nullptr;
    int rc = scratch_address((void **)&base, err);
    if (rc != OK) return rc;
    const size_t o1 = align256(P.s0_bytes), o2 = o1 + align256(P.s1_bytes);
    const Workspace W{P.s0_bytes ? base : nullptr, P.s1_bytes ? base + o1 : nullptr, P.tmp_bytes ? base + o2 : nullptr,
                      P.s0_bytes, P.s1_bytes, P.tmp_bytes};
    return plan_execute_ws(P, in, out, fct, stream, W, err);
}

static int plan_execute_ws(Plan &P, const void *in, void *out, double fct, hipStream_t stream, const Workspace &W,
                           Error &err)
{
    if (P.empty) return OK;
    if (!in || !out) { err = {E_INVALID, "null data pointer"}; return E_INVALID; }
    if (P.r2r) {
        if (in == out && P.sin != P.sout) { err = {E_INVALID, "stride mismatch"}; return E_INVALID; }
        if (((uintptr_t)in | (uintptr_t)out) % 8) {
            err = {E_UNSUPPORTED, "data pointers must be aligned to the element size"};
            return E_UNSUPPORTED;
        }
        return r2r_execute(P, in, out, fct, stream, err);
    }
    if (P.rfp) {
        if (in == out && P.sin != P.sout) { err = {E_INVALID, "stride mismatch"}; return E_INVALID; }
        if (((uintptr_t)in | (uintptr_t)out) % 8) {
            err = {E_UNSUPPORTED, "data pointers must be aligned to the element size"};
            return E_UNSUPPORTED;
        }
        return rfp_execute(P, in, out, fct, stream, err);
    }
    if (in == out && P.kind != J_C2C) {
        err = {E_INVALID, "in-place r2c/c2r is not supported"};
        return E_INVALID;
    }
    if (in == out && P.sin != P.sout) { err = {E_INVALID, "stride mismatch"}; return E_INVALID; }
    const uintptr_t es = P.dtype == 1 ? 4 : 8;
    const uintptr_t ain = (P.kind == J_R2C) ? es : 2 * es, aout = (P.kind == J_C2R) ? es : 2 * es;
    if (((uintptr_t)in % ain) || ((uintptr_t)out % aout)) {
        err = {E_UNSUPPORTED, "data pointers must be aligned to the element size"};
        return E_UNSUPPORTED;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev != P.device) { err = {E_INVALID, "plan was created on another device"}; return E_INVALID; }
    auto resolve = [&](BufRef r) -> void * {
        switch (r) {
            case REF_IN: return const_cast<void *>(in);
            case REF_OUT: return out;
            case REF_S0: return W.s0;
            case REF_S1: return W.s1;
            case REF_TMP: return W.tmp;
            default: return nullptr;
        }
    };
    const bool prof = P.prof_count < P.prof_max;
    hipEvent_t *ev = prof ? &P.prof_ev[(size_t)P.prof_count * P.passes.size() * 2] : nullptr;
    // byte windows of the arrays around their element-0 pointers
    auto window = [&](BufRef r, long long &lo, long long &span) {
        switch (r) {
            case REF_IN: lo = P.in_lo; span = P.in_hi - P.in_lo; break;
            case REF_OUT: lo = P.out_lo; span = P.out_hi - P.out_lo; break;
            case REF_S0: lo = 0; span = (long long)W.b0; break;
            case REF_S1: lo = 0; span = (long long)W.b1; break;
            case REF_TMP: lo = 0; span = (long long)W.bt; break;
            default: lo = 0; span = 0; break;
        }
    };
    for (size_t pi = 0; pi < P.passes.size(); ++pi) {
        auto &ps = P.passes[pi];
        hipError_t e;
        if (prof) (void)hipEventRecord(ev[2 * pi], stream);
        if (ps.kind == Pass::FFT_LDS || ps.kind == Pass::FFT_POW2) {
            const void *srcp = resolve(ps.src);
            FftParams fp = (ps.has_alt && ((uintptr_t)srcp % (ps.fp_alt.src32 ? 8 : 16)) == 0) ? ps.fp_alt : ps.fp;
            fp.src = srcp;
            fp.dst = resolve(ps.dst);
            window(ps.src, fp.src_lo, fp.src_span);
            window(ps.dst, fp.dst_lo, fp.dst_span);
            fp.fct = ps.uses_fct ? fct : 1.0;
            if (ps.kind == Pass::FFT_POW2) {
                e = launch_fft_pow2(fp, ps.map, ps.grid, ps.block, ps.lds, stream);
                ps.kname = "k_fft_pow2<" + std::to_string(fp.N) + ", " + std::to_string(ps.map) + ", " +
                           std::to_string(pow2_io(fp, ps.map, ps.grid, ps.block, ps.lds)) + ">";
            } else {
                e = launch_fft_lds(fp, ps.grid, ps.block, ps.lds, stream);
                ps.kname = "k_fft_lds<" + std::to_string(fft_lds_wpe(fp)) + ">";
            }
        } else {
            RealPostParams rp = ps.rp;
            rp.Z = (const double2 *)resolve(ps.src);
            rp.dst = resolve(ps.dst);
            rp.fct = ps.uses_fct ? fct : 1.0;
            e = launch_real_post(rp, stream);
            ps.kname = "k_real_post";
        }
        if (prof) (void)hipEventRecord(ev[2 * pi + 1], stream);
        if (e != hipSuccess) {
            err = {E_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e)};
            return E_HIP;
        }
    }
    if (prof) ++P.prof_count;
    return OK;
}

int plan_profile(Plan &P, int max_executes, Error &err)
{
    for (auto e : P.prof_ev) (void)hipEventDestroy(e);
    P.prof_ev.clear();
    P.prof_max = P.prof_count = 0;
    if (max_executes <= 0) return OK;
    const size_t n = (size_t)max_executes * P.passes.size() * 2;
    P.prof_ev.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (hipEventCreate(&P.prof_ev[i]) != hipSuccess) {
            (void)hipGetLastError();
            P.prof_ev.resize(i);
            err = {E_HIP, "hipEventCreate failed"};
            return E_HIP;
        }
    }
    P.prof_max = max_executes;
    return OK;
}

int plan_profile_read(Plan &P, double *pass_ms, int npass, int *executes, Error &err)
{
    const size_t np = P.passes.size();
    for (int i = 0; i < npass; ++i) pass_ms[i] = 0.0;
    for (int x = 0; x < P.prof_count; ++x)
        for (size_t pi = 0; pi < np; ++pi) {
            hipEvent_t a = P.prof_ev[((size_t)x * np + pi) * 2], b = P.prof_ev[((size_t)x * np + pi) * 2 + 1];
            float ms = 0.f;
            if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                (void)hipGetLastError();
                err = {E_HIP, "reading profile events failed"};
                return E_HIP;
            }
            if ((int)pi < npass) pass_ms[pi] += ms;
        }
    if (executes) *executes = P.prof_count;
    return OK;
}

// ---------------------------------------------------------------------------
// host staging
// ---------------------------------------------------------------------------

HostStage::~HostStage()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (int b = 0; b < 2; ++b) {
        if (ev[b]) (void)hipEventDestroy(ev[b]);
        if (pin[b]) (void)hipHostFree(pin[b]);
    }
    if (stream) (void)hipStreamDestroy(stream);
}

static int stage_fail(Error &err, hipError_t e, const char *what)
{
    (void)hipGetLastError();
    err = {e == hipErrorOutOfMemory ? E_NOMEM : E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
    return err.code;
}

int HostStage::init(Error &err)
{
    if (stream) return OK;
    // chunk of the pinned ring (IMPULSE_FFT_STAGE_KB, default 4 MiB)
    const size_t ch = (size_t)env_ll("IMPULSE_FFT_STAGE_KB", 4096) * 1024;
    hipError_t e;
    for (int b = 0; b < 2; ++b) {
        if ((e = hipHostMalloc(&pin[b], ch, hipHostMallocDefault)) != hipSuccess) {
            pin[b] = nullptr;
            return stage_fail(err, e, "pinned staging allocation failed");
        }
        if ((e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming)) != hipSuccess) {
            ev[b] = nullptr;
            return stage_fail(err, e, "hipEventCreate failed");
        }
    }
    if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess) {
        stream = nullptr;
        return stage_fail(err, e, "hipStreamCreate failed");
    }
    chunk = ch;
    return OK;
}

int HostStage::to_device(void *dev, const void *host, size_t n, Error &err)
{
    hipError_t e;
    for (size_t off = 0, c = 0; off < n; off += chunk, ++c) {
        const int b = (int)(c & 1);
        const size_t len = std::min(chunk, n - off);
        if (pending[b] && (e = hipEventSynchronize(ev[b])) != hipSuccess) return stage_fail(err, e, "staging copy failed");
        pending[b] = false;
        std::memcpy(pin[b], (const char *)host + off, len);
        if ((e = hipMemcpyAsync((char *)dev + off, pin[b], len, hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipEventRecord(ev[b], stream)) != hipSuccess)
            return stage_fail(err, e, "staging copy failed");
        pending[b] = true;
    }
    return OK;
}

int HostStage::to_host(void *host, const void *dev, size_t n, Error &err)
{
    hipError_t e;
    const size_t nch = (n + chunk - 1) / chunk;
    auto issue = [&](size_t c) -> hipError_t {
        const int b = (int)(c & 1);
        hipError_t r;
        if (pending[b] && (r = hipEventSynchronize(ev[b])) != hipSuccess) return r;
        const size_t off = c * chunk, len = std::min(chunk, n - off);
        if ((r = hipMemcpyAsync(pin[b], (const char *)dev + off, len, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return r;
        if ((r = hipEventRecord(ev[b], stream)) != hipSuccess) return r;
        pending[b] = true;
        return hipSuccess;
    };
    for (size_t c = 0; c < nch && c < 2; ++c)
        if ((e = issue(c)) != hipSuccess) return stage_fail(err, e, "staging copy failed");
    for (size_t c = 0; c < nch; ++c) {
        const int b = (int)(c & 1);
        if ((e = hipEventSynchronize(ev[b])) != hipSuccess) return stage_fail(err, e, "staging copy failed");
        pending[b] = false;
        const size_t off = c * chunk;
        std::memcpy((char *)host + off, pin[b], std::min(chunk, n - off));
        if (c + 2 < nch && (e = issue(c + 2)) != hipSuccess) return stage_fail(err, e, "staging copy failed");
    }
    return OK;
}

int HostStage::sync(Error &err)
{
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return stage_fail(err, e, "stream synchronisation failed");
    pending[0] = pending[1] = false;
    return OK;
}

// The same argument rules as plan_execute (pocketfft_hdronly.h:446-477) are
// applied before anything is staged: in-place calls need equal strides (and
// c2c for the complex kinds), so that the input and output windows coincide and
// one staging buffer with one base serves both.
int plan_execute_host(Plan &P, const void *in, void *out, double fct, Error &err)
{
    if (P.empty) return OK;
    if (!in || !out) { err = {E_INVALID, "null data pointer"}; return E_INVALID; }
    const bool inplace = in == out;
    if (inplace) {
        if (!P.r2r && !P.rfp && P.kind != J_C2C) { err = {E_INVALID, "in-place r2c/c2r is not supported"}; return E_INVALID; }
        if (P.sin != P.sout) { err = {E_INVALID, "stride mismatch"}; return E_INVALID; }
        if (P.in_lo != P.out_lo || P.in_hi != P.out_hi) { err = {E_INVALID, "stride mismatch"}; return E_INVALID; }
    }
    std::lock_guard<std::mutex> lk(P.host_mtx);
    int rc = P.host.init(err);
    if (rc != OK) return rc;
    const size_t nin = (size_t)(P.in_hi - P.in_lo), nout = (size_t)(P.out_hi - P.out_lo);
    if (P.stage_in.bytes < nin && P.stage_in.alloc(nin) != OK) { err = {E_NOMEM, "staging allocation failed"}; return E_NOMEM; }
    if (!inplace && P.stage_out.bytes < nout && P.stage_out.alloc(nout) != OK) {
        err = {E_NOMEM, "staging allocation failed"};
        return E_NOMEM;
    }
    char *din = (char *)P.stage_in.p, *dout = inplace ? din : (char *)P.stage_out.p;
    const char *hin = (const char *)in + P.in_lo;
    char *hout = (char *)out + P.out_lo;
    if ((rc = P.host.to_device(din, hin, nin, err)) != OK) return rc;
    // an output with gaps keeps the bytes the transform does not write
    if (!inplace && !P.out_dense && (rc = P.host.to_device(dout, hout, nout, err)) != OK) return rc;
    if ((rc = plan_execute(P, din - P.in_lo, dout - P.out_lo, fct, P.host.stream, err)) != OK) {
        Error e2;
        (void)P.host.sync(e2);
        return rc;
    }
    return P.host.to_host(hout, dout, nout, err);
}

}  // namespace ifft
