// Host-side engine: plans, device tables and the pass list one execute runs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fft_params.h"

namespace ifft {

// launchers (fft_lds.hip)
hipError_t launch_fft_lds(const FftParams &P, unsigned grid, unsigned block, size_t lds_bytes, hipStream_t s);
int fft_lds_wpe(const FftParams &P);   // waves per SIMD of the generic kernel instantiation a launch takes
hipError_t launch_real_post(const RealPostParams &P, hipStream_t s);
hipError_t launch_pack_halfcomplex(const double2 *bins, double *hc, long long n, hipStream_t s);
hipError_t launch_unpack_halfcomplex(const double *hc, double2 *bins, long long n, hipStream_t s);
// fft_pow2.hip
bool pow2_supported(int N);
int pow2_lds_bytes(int N, int map);
int pow2_tb_host(int N, int map);
int pow2_io(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds);
long long pow2_rowpair_grid(long long signals, long long L2, int TB);
hipError_t launch_fft_pow2(const FftParams &P, int map, unsigned grid, unsigned block, size_t lds, hipStream_t s);

// status codes (mirror impulse_fft_status)
enum Status { OK = 0, E_INVALID = 1, E_NOMEM = 2, E_HIP = 3, E_UNSUPPORTED = 4, E_NODEVICE = 5 };

struct Error {
    int code;
    std::string msg;
};

// A device allocation owned by a plan (freed on destruction).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int alloc(size_t n);
};

// Host <-> device copies of the host-pointer entry points: pinned staging
// buffers used in turn, so that the CPU copy of one chunk overlaps the DMA of
// the previous one, on a stream of its own (default priority).  Not
// thread-safe: the owner serialises calls.
struct HostStage {
    hipStream_t stream = nullptr;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    size_t chunk = 0;
    HostStage() = default;
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    ~HostStage();
    int init(Error &err);
    // host -> device, ordered on `stream`; the host buffer may be reused on return
    int to_device(void *dev, const void *host, size_t n, Error &err);
    // device -> host after the work already queued on `stream`; complete on return
    int to_host(void *host, const void *dev, size_t n, Error &err);
    int sync(Error &err);
};

// Which buffer a pass reads or writes; resolved at execute time.
enum BufRef : int { REF_IN = 0, REF_OUT = 1, REF_S0 = 2, REF_S1 = 3, REF_TMP = 4, REF_NONE = 5 };

struct Pass {
    enum Kind { FFT_LDS, FFT_POW2, REAL_POST } kind;
    int map = 0;                          // FFT_POW2 lane mapping (0 row, 1 col, 2 row load / col store)
    FftParams fp{};
    RealPostParams rp{};
    unsigned grid = 1, block = 64;
    size_t lds = 0;
    BufRef src = REF_NONE, dst = REF_NONE;
    bool uses_fct = false;
    size_t bytes_in = 0, bytes_out = 0;   // bytes the launch must move at least once
    const char *name = "";
    std::string kname;                    // kernel instantiation of the last execute, as rocprof names it
    // even-length real input at unit stride read as complex pairs (used when the
    // source pointer is 16-byte aligned at execute time)
    bool has_alt = false;
    FftParams fp_alt{};
};

struct R2r;   // DCT / DST plan data (dct.hip)
struct Rfp;   // r2r_fftpack / Hartley plan data (hartley.hip)

struct Plan {
    int kind = 0, dtype = 0, forward = 1;
    int device = 0;
    std::vector<size_t> shape;          // shape as given (real shape for r2c/c2r)
    std::vector<ptrdiff_t> sin, sout;   // byte strides
    std::vector<size_t> axes;
    bool empty = false;                 // prod(shape) == 0: execute is a no-op
    std::vector<Pass> passes;
    std::vector<std::shared_ptr<DevBuf>> tables;
    DevBuf s0, s1, tmp;                 // workspace
    size_t s0_bytes = 0, s1_bytes = 0, tmp_bytes = 0;
    // host staging (execute_host)
    std::mutex host_mtx;
    DevBuf stage_in, stage_out;
    HostStage host;
    ptrdiff_t in_lo = 0, in_hi = 0, out_lo = 0, out_hi = 0;   // byte spans relative to element 0
    bool out_dense = false;
    // per-launch event timing (impulse_fft_plan_profile)
    std::vector<hipEvent_t> prof_ev;    // 2 per launch per recorded execute
    int prof_max = 0, prof_count = 0;
    R2r *r2r = nullptr;                 // DCT / DST plans: per-axis pre/post passes around c2c sub-plans
    Rfp *rfp = nullptr;                 // r2r_fftpack / Hartley plans: r2c / c2r sub-plans and a pack pass
    size_t device_bytes() const;
    ~Plan();
};

int plan_create(std::unique_ptr<Plan> &out, int kind, int dtype, size_t ndim, const size_t *shape,
                const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes, const size_t *axes,
                int forward, Error &err);
int plan_execute(Plan &plan, const void *in, void *out, double fct, hipStream_t stream, Error &err);
int plan_execute_host(Plan &plan, const void *in, void *out, double fct, Error &err);
// the same as plan_execute with the workspace in the library's static per-device scratch (no allocation)
int plan_execute_noalloc(Plan &plan, const void *in, void *out, double fct, hipStream_t stream, Error &err);
int plan_profile(Plan &plan, int max_executes, Error &err);
void plan_spans(Plan &P, ptrdiff_t in_elem, ptrdiff_t out_elem);   // host byte spans of the shape/strides

// DCT / DST of types 1-4 (dct.hip); kind 5 = DCT, 6 = DST
int r2r_plan_create(std::unique_ptr<Plan> &out, int kind, int type, int ortho, int dtype, size_t ndim,
                    const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes,
                    const size_t *axes, Error &err);
int r2r_execute(Plan &P, const void *in, void *out, double fct, hipStream_t s, Error &err);
int r2r_num_launches(const R2r &R);
void r2r_destroy(R2r *r);
int plan_profile_read(Plan &plan, double *pass_ms, int npass, int *executes, Error &err);

// r2r_fftpack (kind 7) and the separable (8) / genuine (9) Hartley transforms (hartley.hip)
int rfp_plan_create(std::unique_ptr<Plan> &out, int kind, int r2h, int forward, int dtype, size_t ndim,
                    const size_t *shape, const ptrdiff_t *stride_in, const ptrdiff_t *stride_out, size_t naxes,
                    const size_t *axes, Error &err);
int rfp_execute(Plan &P, const void *in, void *out, double fct, hipStream_t s, Error &err);
int rfp_num_launches(const Rfp &R);
void rfp_destroy(Rfp *r);

// FFT convolution (conv.hip)
struct Conv2d;
int conv2d_create(Conv2d **out, int B, int H, int W, int KH, int KW, int dtype, Error &err);
int conv2d_forward(Conv2d &c, const void *x, const void *k, void *y, hipStream_t s, Error &err);
int conv2d_backward(Conv2d &c, const void *x, const void *k, const void *gy, void *gx, void *gk, hipStream_t s,
                    Error &err);
void conv2d_grid(const Conv2d &c, long long &P, long long &Q);
void conv2d_destroy(Conv2d *c);

// upfirdn: upsample, FIR filter, downsample (upfirdn.hip)
struct Upfirdn;
int upfirdn_create(Upfirdn **out, size_t batch, size_t nt, const double *h, size_t nh, int up, int down, int algo,
                   Error &err);
int upfirdn_execute(Upfirdn &U, const double *t, double *y, hipStream_t s, Error &err);
int upfirdn_execute_host(Upfirdn &U, const double *t, double *y, Error &err);
void upfirdn_info(const Upfirdn &U, size_t *nout, int *algo, size_t *fft_length, size_t *blocks);
void upfirdn_destroy(Upfirdn *u);

// ragged batches of real transforms (ragged.hip)
struct Ragged;
int ragged_create(Ragged **out, size_t count, const size_t *lengths, const size_t *in_off, const size_t *out_off,
                  int forward, Error &err);
int ragged_execute(Ragged &R, const float *x, double2 *y, double fct, hipStream_t s, bool graph, Error &err);
int ragged_tune(Ragged &R, const float *x, double2 *y, hipStream_t s, int reps, double *best_ms, Error &err);
int ragged_configure(Ragged &R, int threads, int lds_max, int order, int waves, Error &err);
int ragged_set_precision(Ragged &R, int precision, Error &err);
void ragged_info(const Ragged &R, int *threads, int *lds_max, int *order, int *waves, int *grid, double *flops,
                 size_t *scratch_bytes);
void ragged_destroy(Ragged *r);

// matrix-core complex FFTs (mfma.hip)
struct Mfma;
int mfma_create(Mfma **out, size_t length, size_t batch, int precision, Error &err);
int mfma_execute(Mfma &m, const void *in, void *out, int forward, double fct, hipStream_t s, Error &err);
size_t mfma_workspace_bytes(const Mfma &m);
void mfma_destroy(Mfma *m);

// host FP64 twiddles (twiddle.hip)
void twiddle_table(std::vector<double2> &out, uint64_t n, uint64_t count, uint64_t step);
double2 twiddle_at(uint64_t k, uint64_t n);   // exp(-2 pi i k / n), exact at multiples of pi/4

// device info
int device_check(Error &err);
int max_lds_bytes();

}  // namespace ifft
