// extern "C" implementation of include/impulse_fft.h.
//
// Every entry point converts internal errors to status codes and records a
// message retrievable with impulse_fft_last_error(); nothing throws across
// the ABI.
#include <cstring>
#include <list>
#include <sstream>
#include <unordered_map>

#include "../../include/impulse_fft.h"
#include "engine.h"

struct impulse_fft_plan_s {
    std::shared_ptr<ifft::Plan> plan;
};

struct impulse_fft_conv2d_s {
    ifft::Conv2d *c = nullptr;
    ~impulse_fft_conv2d_s() { ifft::conv2d_destroy(c); }
};

struct impulse_fft_ragged_s {
    ifft::Ragged *r = nullptr;
    ~impulse_fft_ragged_s() { ifft::ragged_destroy(r); }
};

struct impulse_fft_mfma_s {
    ifft::Mfma *m = nullptr;
    ~impulse_fft_mfma_s() { ifft::mfma_destroy(m); }
};

struct impulse_fft_upfirdn_s {
    ifft::Upfirdn *u = nullptr;
    ~impulse_fft_upfirdn_s() { ifft::upfirdn_destroy(u); }
};

// C path: the device plans behind make_cfft_plan / make_rfft_plan come from
// the plan caches below, so that impulse's plan-per-call pattern
// (c_pocketfft/pocketfft.nim:216-303 builds and destroys a plan in every
// fft() call) pays a cache lookup, not a device allocation and a twiddle
// upload, after the first call of a length.
struct impulse_cfft_plan_s {
    size_t length = 0;
    std::shared_ptr<ifft::Plan> fwd, bwd;
};

namespace {
// one real length of the C path: r2c / c2r plans and the device buffers of the
// halfcomplex repacking, shared by every rfft plan handle of that length
struct RfftCore {
    std::shared_ptr<ifft::Plan> r2c, c2r;
    ifft::DevBuf x, bins, hc;   // device staging: reals, complex bins, halfcomplex
    ifft::HostStage st;
    std::mutex mtx;
};
}  // namespace

struct impulse_rfft_plan_s {
    size_t length = 0;
    std::shared_ptr<RfftCore> core;
};

namespace {

thread_local std::string g_last_error;

int fail(const ifft::Error &e)
{
    g_last_error = e.msg;
    return e.code ? e.code : ifft::E_HIP;
}

int fail(int code, const char *msg)
{
    g_last_error = msg;
    return code;
}

// ---- LRU plan caches ---------------------------------------------------------
// Like get_plan (pocketfft_hdronly.h:2655-2706): a mutex-guarded LRU list,
// looked up under the lock, built outside it, and inserted with a second
// lookup in case another thread built the same plan meanwhile.
template <class V> struct Lru {
    std::mutex mtx;
    size_t capacity = 16;
    size_t hits = 0, misses = 0;
    std::list<std::pair<std::string, std::shared_ptr<V>>> lru;
    std::unordered_map<std::string, typename decltype(lru)::iterator> map;

    void trim()
    {
        while (lru.size() > capacity) {
            map.erase(lru.back().first);
            lru.pop_back();
        }
    }
    // most recently used entry of `key`, or null (counts a hit or a miss)
    std::shared_ptr<V> find(const std::string &key)
    {
        std::lock_guard<std::mutex> lk(mtx);
        auto it = map.find(key);
        if (it == map.end()) {
            ++misses;
            return nullptr;
        }
        ++hits;
        lru.splice(lru.begin(), lru, it->second);
        return it->second->second;
    }
    // insert `v` unless another thread inserted `key` first; returns the entry kept
    std::shared_ptr<V> insert(const std::string &key, std::shared_ptr<V> v)
    {
        std::lock_guard<std::mutex> lk(mtx);
        auto it = map.find(key);
        if (it != map.end()) return it->second->second;
        lru.emplace_front(key, v);
        map[key] = lru.begin();
        trim();
        return v;
    }
};

Lru<ifft::Plan> &cache()
{
    static auto *c = new Lru<ifft::Plan>();   // never destroyed: plans may outlive static teardown
    return *c;
}

Lru<RfftCore> &rcache()
{
    static auto *c = new Lru<RfftCore>();
    return *c;
}

std::string cache_key(int kind, int dtype, size_t ndim, const size_t *shape, const ptrdiff_t *sin,
                      const ptrdiff_t *sout, size_t naxes, const size_t *axes, int forward, int dev)
{
    std::ostringstream k;
    // `forward` is a flag for the FFT kinds and 2 type + ortho for the DCT / DST kinds
    k << kind << '|' << dtype << '|' << (kind >= IMPULSE_FFT_DCT ? forward : (forward ? 1 : 0)) << '|' << dev << '|';
    for (size_t d = 0; d < ndim; ++d) k << shape[d] << ',' << sin[d] << ',' << sout[d] << ';';
    k << '|';
    for (size_t a = 0; a < naxes; ++a) k << axes[a] << ',';
    return k.str();
}

int cached_plan(std::shared_ptr<ifft::Plan> &out, int kind, int dtype, size_t ndim, const size_t *shape,
                const ptrdiff_t *sin, const ptrdiff_t *sout, size_t naxes, const size_t *axes, int forward)
{
    if (ndim < 1 || !shape || !sin || !sout || naxes < 1 || !axes)
        return fail(ifft::E_INVALID, "invalid argument");
    int dev = 0;
    (void)hipGetDevice(&dev);
    const std::string key = cache_key(kind, dtype, ndim, shape, sin, sout, naxes, axes, forward, dev);
    auto &c = cache();
    if ((out = c.find(key))) return ifft::OK;
    std::unique_ptr<ifft::Plan> p;
    ifft::Error err{0, ""};
    // DCT / DST (kinds 5, 6) carry the type and ortho flag in `forward` (2 type + ortho), the
    // FFTPACK / Hartley kinds (7-9) real2hermitian and the direction (2 real2hermitian + forward)
    int rc = kind >= 7   ? ifft::rfp_plan_create(p, kind, forward >> 1, forward & 1, dtype, ndim, shape, sin, sout,
                                                 naxes, axes, err)
             : kind >= 5 ? ifft::r2r_plan_create(p, kind, forward >> 1, forward & 1, dtype, ndim, shape, sin, sout,
                                                 naxes, axes, err)
                         : ifft::plan_create(p, kind, dtype, ndim, shape, sin, sout, naxes, axes, forward, err);
    if (rc != ifft::OK) return fail(err);
    out = c.insert(key, std::shared_ptr<ifft::Plan>(std::move(p)));
    return ifft::OK;
}

int oneshot(int kind, size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
            size_t naxes, const size_t *axes, int forward, int dtype, const void *in, void *out, double fct)
{
    if (shape)
        for (size_t d = 0; d < ndim; ++d)
            if (shape[d] == 0) return ifft::OK;   // pocketfft returns early on empty arrays
    // in-place rules before any plan is built (pocketfft_hdronly.h:453-454): equal
    // strides, and no in-place r2c / c2r
    if (in && in == out) {
        if (kind == IMPULSE_FFT_R2C || kind == IMPULSE_FFT_C2R || kind == IMPULSE_FFT_R2C_ADJOINT ||
            kind == IMPULSE_FFT_C2R_ADJOINT)
            return fail(ifft::E_INVALID, "in-place r2c/c2r is not supported");
        if (sin && sout)
            for (size_t d = 0; d < ndim; ++d)
                if (sin[d] != sout[d]) return fail(ifft::E_INVALID, "stride mismatch");
    }
    std::shared_ptr<ifft::Plan> p;
    int rc = cached_plan(p, kind, dtype, ndim, shape, sin, sout, naxes, axes, forward);
    if (rc != ifft::OK) return rc;
    ifft::Error err{0, ""};
    rc = ifft::plan_execute_host(*p, in, out, fct, err);
    return rc == ifft::OK ? rc : fail(err);
}

// 1-D FP64 plan of the C path, unit stride, from the plan cache
int cached_1d(std::shared_ptr<ifft::Plan> &p, int kind, size_t n, int forward)
{
    const size_t shape[1] = {n};
    const ptrdiff_t sin[1] = {kind == IMPULSE_FFT_R2C ? 8 : 16};
    const ptrdiff_t sout[1] = {kind == IMPULSE_FFT_C2R ? 8 : 16};
    const size_t axes[1] = {0};
    return cached_plan(p, kind, IMPULSE_FFT_F64, 1, shape, sin, sout, 1, axes, forward);
}

}  // namespace

extern "C" {

int impulse_fft_plan_create(impulse_fft_plan *plan, int kind, int dtype, size_t ndim, const size_t *shape,
                            const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                            int forward)
{
    if (!plan) return fail(ifft::E_INVALID, "plan pointer is NULL");
    *plan = nullptr;
    std::unique_ptr<ifft::Plan> p;
    ifft::Error err{0, ""};
    int rc = ifft::plan_create(p, kind, dtype, ndim, shape, stride_in, stride_out, naxes, axes, forward, err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_plan_s;
    if (!h) return fail(ifft::E_NOMEM, "out of host memory");
    h->plan = std::move(p);
    *plan = h;
    return ifft::OK;
}

int impulse_fft_execute(impulse_fft_plan plan, const void *data_in, void *data_out, double fct, void *stream)
{
    if (!plan || !plan->plan) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::plan_execute(*plan->plan, data_in, data_out, fct, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_execute_noalloc(impulse_fft_plan plan, const void *data_in, void *data_out, double fct, void *stream)
{
    if (!plan || !plan->plan) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::plan_execute_noalloc(*plan->plan, data_in, data_out, fct, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_execute_host(impulse_fft_plan plan, const void *data_in, void *data_out, double fct)
{
    if (!plan || !plan->plan) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::plan_execute_host(*plan->plan, data_in, data_out, fct, err);
    return rc == ifft::OK ? rc : fail(err);
}

size_t impulse_fft_plan_device_bytes(impulse_fft_plan plan)
{
    return (plan && plan->plan) ? plan->plan->device_bytes() : 0;
}

int impulse_fft_plan_num_launches(impulse_fft_plan plan)
{
    if (!plan || !plan->plan) return -1;
    if (plan->plan->r2r) return ifft::r2r_num_launches(*plan->plan->r2r);
    if (plan->plan->rfp) return ifft::rfp_num_launches(*plan->plan->rfp);
    return (int)plan->plan->passes.size();
}

int impulse_fft_plan_pass_info(impulse_fft_plan plan, int pass, char *name, size_t name_len, size_t *bytes_in,
                               size_t *bytes_out, unsigned *grid, unsigned *block, size_t *lds_bytes)
{
    if (!plan || !plan->plan) return fail(ifft::E_INVALID, "plan is NULL");
    if (plan->plan->r2r || plan->plan->rfp)
        return fail(ifft::E_UNSUPPORTED, "per-launch information is not kept for real-to-real plans");
    const auto &P = *plan->plan;
    if (pass < 0 || pass >= (int)P.passes.size()) return fail(ifft::E_INVALID, "pass index out of range");
    const auto &ps = P.passes[pass];
    if (name && name_len) {
        std::strncpy(name, ps.kname.empty() ? ps.name : ps.kname.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (bytes_in) *bytes_in = ps.bytes_in;
    if (bytes_out) *bytes_out = ps.bytes_out;
    if (grid) *grid = ps.kind != ifft::Pass::REAL_POST ? ps.grid : 0;
    if (block) *block = ps.kind != ifft::Pass::REAL_POST ? ps.block : 256;
    if (lds_bytes) *lds_bytes = ps.lds;
    return ifft::OK;
}

int impulse_fft_plan_profile(impulse_fft_plan plan, int max_executes)
{
    if (!plan || !plan->plan) return fail(ifft::E_INVALID, "plan is NULL");
    if (plan->plan->r2r || plan->plan->rfp)
        return fail(ifft::E_UNSUPPORTED, "per-launch timing is not kept for real-to-real plans");
    ifft::Error err{0, ""};
    int rc = ifft::plan_profile(*plan->plan, max_executes, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_plan_profile_read(impulse_fft_plan plan, double *pass_ms, int npass, int *executes)
{
    if (!plan || !plan->plan || (!pass_ms && npass > 0)) return fail(ifft::E_INVALID, "invalid argument");
    if (plan->plan->r2r || plan->plan->rfp)
        return fail(ifft::E_UNSUPPORTED, "per-launch timing is not kept for real-to-real plans");
    ifft::Error err{0, ""};
    int rc = ifft::plan_profile_read(*plan->plan, pass_ms, npass, executes, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_plan_destroy(impulse_fft_plan plan)
{
    delete plan;
    return ifft::OK;
}

int impulse_fft_c2c(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                    size_t naxes, const size_t *axes, int forward, int dtype, const void *data_in, void *data_out,
                    double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_C2C, ndim, shape, stride_in, stride_out, naxes, axes, forward, dtype, data_in,
                   data_out, fct);
}

int impulse_fft_r2c(size_t ndim, const size_t *shape_in, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                    size_t naxes, const size_t *axes, int forward, int dtype, const void *data_in, void *data_out,
                    double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_R2C, ndim, shape_in, stride_in, stride_out, naxes, axes, forward, dtype, data_in,
                   data_out, fct);
}

int impulse_fft_r2r_plan_create(impulse_fft_plan *plan, int kind, int type, int ortho, int dtype, size_t ndim,
                                const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                                size_t naxes, const size_t *axes)
{
    if (!plan) return fail(ifft::E_INVALID, "plan pointer is NULL");
    *plan = nullptr;
    std::unique_ptr<ifft::Plan> p;
    ifft::Error err{0, ""};
    int rc = ifft::r2r_plan_create(p, kind, type, ortho, dtype, ndim, shape, stride_in, stride_out, naxes, axes, err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_plan_s;
    if (!h) return fail(ifft::E_NOMEM, "out of host memory");
    h->plan = std::move(p);
    *plan = h;
    return ifft::OK;
}

static int oneshot_r2r(int kind, size_t ndim, const size_t *shape, const ptrdiff_t *sin, const ptrdiff_t *sout,
                       size_t naxes, const size_t *axes, int type, int dtype, const void *in, void *out, double fct,
                       int ortho)
{
    if (type < 1 || type > 4) return fail(ifft::E_INVALID, kind == 5 ? "invalid DCT type" : "invalid DST type");
    return oneshot(kind, ndim, shape, sin, sout, naxes, axes, 2 * type + (ortho ? 1 : 0), dtype, in, out, fct);
}

int impulse_fft_fftpack_plan_create(impulse_fft_plan *plan, int kind, int real2hermitian, int forward, int dtype,
                                    size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes)
{
    if (!plan) return fail(ifft::E_INVALID, "plan pointer is NULL");
    *plan = nullptr;
    std::unique_ptr<ifft::Plan> p;
    ifft::Error err{0, ""};
    int rc = ifft::rfp_plan_create(p, kind, real2hermitian ? 1 : 0, forward ? 1 : 0, dtype, ndim, shape, stride_in,
                                   stride_out, naxes, axes, err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_plan_s;
    if (!h) return fail(ifft::E_NOMEM, "out of host memory");
    h->plan = std::move(p);
    *plan = h;
    return ifft::OK;
}

int impulse_fft_r2r_fftpack(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                            size_t naxes, const size_t *axes, int real2hermitian, int forward, int dtype,
                            const void *data_in, void *data_out, double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_FFTPACK, ndim, shape, stride_in, stride_out, naxes, axes,
                   2 * (real2hermitian ? 1 : 0) + (forward ? 1 : 0), dtype, data_in, data_out, fct);
}

int impulse_fft_r2r_separable_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                      const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int dtype,
                                      const void *data_in, void *data_out, double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_HARTLEY, ndim, shape, stride_in, stride_out, naxes, axes, 0, dtype, data_in,
                   data_out, fct);
}

int impulse_fft_r2r_genuine_hartley(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in,
                                    const ptrdiff_t *stride_out, size_t naxes, const size_t *axes, int dtype,
                                    const void *data_in, void *data_out, double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_HARTLEY_GENUINE, ndim, shape, stride_in, stride_out, naxes, axes, 0, dtype, data_in,
                   data_out, fct);
}

int impulse_fft_dct(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                    size_t naxes, const size_t *axes, int type, int dtype, const void *data_in, void *data_out,
                    double fct, int ortho, size_t nthreads)
{
    (void)nthreads;
    return oneshot_r2r(IMPULSE_FFT_DCT, ndim, shape, stride_in, stride_out, naxes, axes, type, dtype, data_in,
                       data_out, fct, ortho);
}

int impulse_fft_dst(size_t ndim, const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                    size_t naxes, const size_t *axes, int type, int dtype, const void *data_in, void *data_out,
                    double fct, int ortho, size_t nthreads)
{
    (void)nthreads;
    return oneshot_r2r(IMPULSE_FFT_DST, ndim, shape, stride_in, stride_out, naxes, axes, type, dtype, data_in,
                       data_out, fct, ortho);
}

int impulse_fft_c2r(size_t ndim, const size_t *shape_out, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out,
                    size_t naxes, const size_t *axes, int forward, int dtype, const void *data_in, void *data_out,
                    double fct, size_t nthreads)
{
    (void)nthreads;
    return oneshot(IMPULSE_FFT_C2R, ndim, shape_out, stride_in, stride_out, naxes, axes, forward, dtype, data_in,
                   data_out, fct);
}

int impulse_fft_cache_set_capacity(size_t entries)
{
    for (int which = 0; which < 2; ++which) {
        auto apply = [&](auto &c) {
            std::lock_guard<std::mutex> lk(c.mtx);
            c.capacity = entries;
            c.trim();
        };
        if (which == 0) apply(cache());
        else apply(rcache());
    }
    return ifft::OK;
}

size_t impulse_fft_cache_size(void)
{
    auto &c = cache();
    std::lock_guard<std::mutex> lk(c.mtx);
    return c.lru.size();
}

int impulse_fft_cache_clear(void)
{
    {
        auto &c = cache();
        std::lock_guard<std::mutex> lk(c.mtx);
        c.lru.clear();
        c.map.clear();
        c.hits = c.misses = 0;
    }
    auto &r = rcache();
    std::lock_guard<std::mutex> lk(r.mtx);
    r.lru.clear();
    r.map.clear();
    r.hits = r.misses = 0;
    return ifft::OK;
}

int impulse_fft_cache_stats(size_t *hits, size_t *misses)
{
    auto &c = cache();
    std::lock_guard<std::mutex> lk(c.mtx);
    if (hits) *hits = c.hits;
    if (misses) *misses = c.misses;
    return ifft::OK;
}

// ---- upfirdn --------------------------------------------------------------------

int impulse_fft_upfirdn_create(impulse_fft_upfirdn *plan, size_t batch, size_t nt, const double *h, size_t nh,
                               int up, int down, int algorithm)
{
    if (!plan) return fail(ifft::E_INVALID, "plan pointer is NULL");
    *plan = nullptr;
    ifft::Error err{0, ""};
    ifft::Upfirdn *u = nullptr;
    int rc = ifft::upfirdn_create(&u, batch, nt, h, nh, up, down, algorithm, err);
    if (rc != ifft::OK) return fail(err);
    auto *p = new (std::nothrow) impulse_fft_upfirdn_s;
    if (!p) {
        ifft::upfirdn_destroy(u);
        return fail(ifft::E_NOMEM, "out of host memory");
    }
    p->u = u;
    *plan = p;
    return ifft::OK;
}

int impulse_fft_upfirdn_info(impulse_fft_upfirdn plan, size_t *nout, int *algorithm, size_t *fft_length,
                             size_t *blocks)
{
    if (!plan || !plan->u) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::upfirdn_info(*plan->u, nout, algorithm, fft_length, blocks);
    return ifft::OK;
}

int impulse_fft_upfirdn_execute(impulse_fft_upfirdn plan, const double *t, double *y, void *stream)
{
    if (!plan || !plan->u) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::upfirdn_execute(*plan->u, t, y, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_upfirdn_execute_host(impulse_fft_upfirdn plan, const double *t, double *y)
{
    if (!plan || !plan->u) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::upfirdn_execute_host(*plan->u, t, y, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_upfirdn_destroy(impulse_fft_upfirdn plan)
{
    delete plan;
    return ifft::OK;
}

// ---- FFT convolution ----------------------------------------------------------

int impulse_fft_conv2d_create(impulse_fft_conv2d *conv, int batch, int height, int width, int kh, int kw, int dtype)
{
    if (!conv) return fail(ifft::E_INVALID, "conv pointer is NULL");
    *conv = nullptr;
    ifft::Error err{0, ""};
    int rc = ifft::device_check(err);
    if (rc != ifft::OK) return fail(err);
    ifft::Conv2d *c = nullptr;
    rc = ifft::conv2d_create(&c, batch, height, width, kh, kw, dtype == IMPULSE_FFT_F32 ? 1 : (dtype == 0 ? 0 : -1), err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_conv2d_s;
    if (!h) {
        ifft::conv2d_destroy(c);
        return fail(ifft::E_NOMEM, "out of host memory");
    }
    h->c = c;
    *conv = h;
    return ifft::OK;
}

int impulse_fft_conv2d_forward(impulse_fft_conv2d conv, const void *x, const void *k, void *y, void *stream)
{
    if (!conv || !conv->c || !x || !k || !y) return fail(ifft::E_INVALID, "NULL argument");
    ifft::Error err{0, ""};
    int rc = ifft::conv2d_forward(*conv->c, x, k, y, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_conv2d_backward(impulse_fft_conv2d conv, const void *x, const void *k, const void *gy, void *gx,
                                void *gk, void *stream)
{
    if (!conv || !conv->c || !x || !k || !gy) return fail(ifft::E_INVALID, "NULL argument");
    ifft::Error err{0, ""};
    int rc = ifft::conv2d_backward(*conv->c, x, k, gy, gx, gk, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_conv2d_grid(impulse_fft_conv2d conv, size_t *P, size_t *Q)
{
    if (!conv || !conv->c) return fail(ifft::E_INVALID, "conv is NULL");
    long long p, q;
    ifft::conv2d_grid(*conv->c, p, q);
    if (P) *P = (size_t)p;
    if (Q) *Q = (size_t)q;
    return ifft::OK;
}

int impulse_fft_conv2d_destroy(impulse_fft_conv2d conv)
{
    delete conv;
    return ifft::OK;
}

// ---- ragged batches -----------------------------------------------------------

int impulse_fft_ragged_create(impulse_fft_ragged *batch, size_t count, const size_t *lengths,
                              const size_t *in_offsets, const size_t *out_offsets, int forward)
{
    if (!batch) return fail(ifft::E_INVALID, "batch pointer is NULL");
    *batch = nullptr;
    ifft::Error err{0, ""};
    ifft::Ragged *r = nullptr;
    int rc = ifft::ragged_create(&r, count, lengths, in_offsets, out_offsets, forward, err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_ragged_s;
    if (!h) {
        ifft::ragged_destroy(r);
        return fail(ifft::E_NOMEM, "out of host memory");
    }
    h->r = r;
    *batch = h;
    return ifft::OK;
}

int impulse_fft_ragged_execute(impulse_fft_ragged batch, const float *in, void *out, double fct, void *stream)
{
    if (!batch || !batch->r) return fail(ifft::E_INVALID, "batch is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::ragged_execute(*batch->r, in, (double2 *)out, fct, (hipStream_t)stream, true, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_ragged_tune(impulse_fft_ragged batch, const float *in, void *out, void *stream, int reps,
                            double *best_ms)
{
    if (!batch || !batch->r) return fail(ifft::E_INVALID, "batch is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::ragged_tune(*batch->r, in, (double2 *)out, (hipStream_t)stream, reps, best_ms, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_ragged_configure(impulse_fft_ragged batch, int threads, int lds_points, int radix_order,
                                 int waves_per_simd)
{
    if (!batch || !batch->r) return fail(ifft::E_INVALID, "batch is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::ragged_configure(*batch->r, threads, lds_points, radix_order, waves_per_simd, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_ragged_set_precision(impulse_fft_ragged batch, int precision)
{
    if (!batch || !batch->r) return fail(ifft::E_INVALID, "batch is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::ragged_set_precision(*batch->r, precision, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_ragged_info(impulse_fft_ragged batch, int *threads, int *lds_points, int *radix_order,
                            int *waves_per_simd, int *grid, double *flops, size_t *scratch_bytes)
{
    if (!batch || !batch->r) return fail(ifft::E_INVALID, "batch is NULL");
    ifft::ragged_info(*batch->r, threads, lds_points, radix_order, waves_per_simd, grid, flops, scratch_bytes);
    return ifft::OK;
}

int impulse_fft_ragged_destroy(impulse_fft_ragged batch)
{
    delete batch;
    return ifft::OK;
}

// ---- matrix-core FFTs ---------------------------------------------------------

int impulse_fft_mfma_create(impulse_fft_mfma *plan, size_t length, size_t batch, int precision)
{
    if (!plan) return fail(ifft::E_INVALID, "plan pointer is NULL");
    *plan = nullptr;
    ifft::Error err{0, ""};
    ifft::Mfma *m = nullptr;
    int rc = ifft::mfma_create(&m, length, batch, precision, err);
    if (rc != ifft::OK) return fail(err);
    auto *h = new (std::nothrow) impulse_fft_mfma_s;
    if (!h) {
        ifft::mfma_destroy(m);
        return fail(ifft::E_NOMEM, "out of host memory");
    }
    h->m = m;
    *plan = h;
    return ifft::OK;
}

int impulse_fft_mfma_execute(impulse_fft_mfma plan, const void *in, void *out, int forward, double fct, void *stream)
{
    if (!plan || !plan->m) return fail(ifft::E_INVALID, "plan is NULL");
    ifft::Error err{0, ""};
    int rc = ifft::mfma_execute(*plan->m, in, out, forward, fct, (hipStream_t)stream, err);
    return rc == ifft::OK ? rc : fail(err);
}

int impulse_fft_mfma_workspace_bytes(impulse_fft_mfma plan, size_t *bytes)
{
    if (!plan || !plan->m || !bytes) return fail(ifft::E_INVALID, "NULL argument");
    *bytes = ifft::mfma_workspace_bytes(*plan->m);
    return ifft::OK;
}

int impulse_fft_mfma_destroy(impulse_fft_mfma plan)
{
    delete plan;
    return ifft::OK;
}

// ---- C path -----------------------------------------------------------------

impulse_cfft_plan impulse_make_cfft_plan(size_t length)
{
    if (length == 0) { fail(ifft::E_INVALID, "length must be >= 1"); return nullptr; }
    auto *h = new (std::nothrow) impulse_cfft_plan_s;
    if (!h) return nullptr;
    h->length = length;
    if (cached_1d(h->fwd, IMPULSE_FFT_C2C, length, 1) != ifft::OK ||
        cached_1d(h->bwd, IMPULSE_FFT_C2C, length, 0) != ifft::OK) {
        delete h;
        return nullptr;
    }
    return h;
}

void impulse_destroy_cfft_plan(impulse_cfft_plan plan) { delete plan; }

static int cfft_run(impulse_cfft_plan plan, double c[], double fct, bool forward)
{
    if (!plan || !c) { fail(ifft::E_INVALID, "NULL plan or data"); return -1; }
    if (plan->length == 1) return 0;   // PocketFFT's C path returns unscaled here (pocketfft.c:874)
    ifft::Error err{0, ""};
    int rc = ifft::plan_execute_host(forward ? *plan->fwd : *plan->bwd, c, c, fct, err);
    if (rc != ifft::OK) { fail(err); return -1; }
    return 0;
}

int impulse_cfft_forward(impulse_cfft_plan plan, double c[], double fct) { return cfft_run(plan, c, fct, true); }
int impulse_cfft_backward(impulse_cfft_plan plan, double c[], double fct) { return cfft_run(plan, c, fct, false); }
size_t impulse_cfft_length(impulse_cfft_plan plan) { return plan ? plan->length : 0; }

impulse_rfft_plan impulse_make_rfft_plan(size_t length)
{
    if (length == 0) { fail(ifft::E_INVALID, "length must be >= 1"); return nullptr; }
    auto *h = new (std::nothrow) impulse_rfft_plan_s;
    if (!h) return nullptr;
    h->length = length;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const std::string key = std::to_string(length) + '|' + std::to_string(dev);
    auto &rc = rcache();
    if (!(h->core = rc.find(key))) {
        auto core = std::make_shared<RfftCore>();
        const size_t nb = length / 2 + 1;
        if (cached_1d(core->r2c, IMPULSE_FFT_R2C, length, 1) != ifft::OK ||
            cached_1d(core->c2r, IMPULSE_FFT_C2R, length, 0) != ifft::OK || core->x.alloc(length * 8) != ifft::OK ||
            core->bins.alloc(nb * 16) != ifft::OK || core->hc.alloc(length * 8) != ifft::OK) {
            if (g_last_error.empty()) g_last_error = "rfft plan allocation failed";
            delete h;
            return nullptr;
        }
        h->core = rc.insert(key, core);
    }
    return h;
}

void impulse_destroy_rfft_plan(impulse_rfft_plan plan) { delete plan; }

static int rfft_run(impulse_rfft_plan plan, double c[], double fct, bool forward)
{
    if (!plan || !c) { fail(ifft::E_INVALID, "NULL plan or data"); return -1; }
    const size_t n = plan->length;
    if (n == 1) return 0;   // PocketFFT's C path returns unscaled here (pocketfft.c:1703,1739)
    RfftCore &C = *plan->core;
    std::lock_guard<std::mutex> lk(C.mtx);
    ifft::Error err{0, ""};
    int rc = C.st.init(err);
    hipError_t e = hipSuccess;
    if (rc == ifft::OK && forward) {
        rc = C.st.to_device(C.x.p, c, n * 8, err);
        if (rc == ifft::OK) rc = ifft::plan_execute(*C.r2c, C.x.p, C.bins.p, fct, C.st.stream, err);
        if (rc == ifft::OK)
            e = ifft::launch_pack_halfcomplex((const double2 *)C.bins.p, (double *)C.hc.p, (long long)n, C.st.stream);
        if (rc == ifft::OK && e == hipSuccess) rc = C.st.to_host(c, C.hc.p, n * 8, err);
    } else if (rc == ifft::OK) {
        rc = C.st.to_device(C.hc.p, c, n * 8, err);
        if (rc == ifft::OK)
            e = ifft::launch_unpack_halfcomplex((const double *)C.hc.p, (double2 *)C.bins.p, (long long)n, C.st.stream);
        if (rc == ifft::OK && e == hipSuccess) rc = ifft::plan_execute(*C.c2r, C.bins.p, C.x.p, fct, C.st.stream, err);
        if (rc == ifft::OK && e == hipSuccess) rc = C.st.to_host(c, C.x.p, n * 8, err);
    }
    if (rc != ifft::OK || e != hipSuccess) {
        ifft::Error e2;
        if (C.st.stream) (void)C.st.sync(e2);
        if (rc != ifft::OK) fail(err);
        else {
            (void)hipGetLastError();
            fail(ifft::E_HIP, hipGetErrorString(e));
        }
        return -1;
    }
    return 0;
}

int impulse_rfft_forward(impulse_rfft_plan plan, double c[], double fct) { return rfft_run(plan, c, fct, true); }
int impulse_rfft_backward(impulse_rfft_plan plan, double c[], double fct) { return rfft_run(plan, c, fct, false); }
size_t impulse_rfft_length(impulse_rfft_plan plan) { return plan ? plan->length : 0; }

// ---- diagnostics ------------------------------------------------------------

const char *impulse_fft_status_string(int status)
{
    switch (status) {
        case IMPULSE_FFT_OK: return "ok";
        case IMPULSE_FFT_ERR_INVALID_ARG: return "invalid argument";
        case IMPULSE_FFT_ERR_NOMEM: return "out of memory";
        case IMPULSE_FFT_ERR_HIP: return "HIP runtime error";
        case IMPULSE_FFT_ERR_UNSUPPORTED: return "unsupported request";
        case IMPULSE_FFT_ERR_NO_DEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

const char *impulse_fft_last_error(void) { return g_last_error.c_str(); }

int impulse_fft_device_check(void)
{
    ifft::Error err{0, ""};
    int rc = ifft::device_check(err);
    return rc == ifft::OK ? rc : fail(err);
}

}  // extern "C"
