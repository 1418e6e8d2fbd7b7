// cts_surrogate.cpp -- C ABI of the learned field surrogate
// (include/cetasim_hip.h, "field surrogate").  The reference has no
// counterpart: it evaluates every Bassetti-Erskine kick with the Faddeeva
// function (src/Bunch.cpp:119-211, src/MPBunch.cpp:1094-1210).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cts_internal.hpp"
#include "cts_mlp.hpp"

using namespace cts;

namespace cts {
hipError_t launch_mlp_samples(int64_t n, uint64_t seed, uint32_t tag, float *out, double *exact, hipStream_t st);
hipError_t launch_mlp_train(const float *params, const float *samples, int64_t n, float *part, const float norm[4],
                            float inv_batch, uint64_t seed, uint32_t step, int chunks, int random, int nwg,
                            hipStream_t st);
hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *part, int nwg, float lr, float b1,
                           float b2, float eps, float bc1, float bc2, float *grad_out, float *loss_out,
                           hipStream_t st);
hipError_t launch_mlp_eval(const float *W, const float *feat, int64_t n, int stride, float *out, hipStream_t st);
hipError_t launch_mlp_field_error(const float *W, const double *exact, int64_t n, const float norm[4], double *err,
                                  hipStream_t st);
hipError_t upload_faddeeva_tables_mlp(const FadTables &t);
}  // namespace cts

// Staging of cts_sur_eval_noalloc: features (3 floats) and outputs (2
// floats) of up to kSurScratchPoints points, 5 MiB per device, held until
// the process ends.  In host code the array's name is not a device address:
// the kernels and copies take the per-device pointer that
// hipGetSymbolAddress returns (scratch_ptr).
constexpr int64_t kSurScratchPoints = 1 << 18;
static __device__ __align__(256) float scratch[kSurScratchPoints * 5];

struct cts_surrogate {
    hipStream_t stream = nullptr;
    float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr, *d_part = nullptr, *d_grad = nullptr;
    float *d_loss = nullptr;
    float *d_samples = nullptr;  // [n_train][5]
    double *d_val = nullptr;     // [n_val][5] exact |u| |v| r Gx Gy
    double *d_err = nullptr;     // [n_val]
    int64_t n_train = 0, n_val = 0;
    float norm[4] = {0, 1, 0, 1};
    int part_cap = 0;            // workgroups the partial buffer holds
    int64_t step = 0;            // Adam steps taken
};

namespace {

thread_local std::string s_err;

int sfail(int code, const std::string &msg)
{
    s_err = msg;
    return code;
}

#define SCHK(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return sfail(CTS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int salloc(T **p, size_t n)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) e = hipMemset(*p, 0, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) return sfail(CTS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return CTS_OK;
}

int ensure_part(cts_surrogate *s, int nwg)
{
    if (nwg <= s->part_cap) return CTS_OK;
    int rc = salloc(&s->d_part, (size_t)nwg * (kMlpParams + 1));
    if (rc) return rc;
    s->part_cap = nwg;
    return CTS_OK;
}

// splitmix64 stream for the host-side initialisation
uint64_t mix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// scratch's device address on each device, looked up once per device
// with that device current
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};

int scratch_ptr(float **out)
{
    int dev = 0;
    SCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return sfail(CTS_ERR_STATE, "scratch: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        SCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (float *)p;
    }
    *out = g_scratch[dev];
    return CTS_OK;
}

}  // namespace

namespace cts {
// device parameters and output normalisation for the kick kernels
void sur_export(const cts_surrogate *s, const float **params, SurNorm *norm)
{
    *params = s->d_params;
    norm->mx = s->norm[0];
    norm->sx = s->norm[1];
    norm->my = s->norm[2];
    norm->sy = s->norm[3];
}
}  // namespace cts

extern "C" {

const char *cts_sur_last_error(void) { return s_err.c_str(); }

int cts_sur_create(cts_surrogate **out)
{
    if (!out) return sfail(CTS_ERR_ARG, "cts_sur_create: null");
    cts_surrogate *s = new cts_surrogate();
    hipError_t e = hipStreamCreate(&s->stream);
    if (e != hipSuccess) {
        delete s;
        return sfail(CTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    }
    // w(z) tables of the sample generator's FP64 field
    FadTables ft;
    build_fad_tables(ft);
    int rc;
    if ((e = upload_faddeeva_tables_mlp(ft)) != hipSuccess ||
        (rc = salloc(&s->d_params, kMlpParams)) || (rc = salloc(&s->d_m, kMlpParams)) ||
        (rc = salloc(&s->d_v, kMlpParams)) || (rc = salloc(&s->d_grad, kMlpParams)) || (rc = salloc(&s->d_loss, 1))) {
        if (e != hipSuccess) sfail(CTS_ERR_HIP, std::string("tables: ") + hipGetErrorString(e));
        (void)hipStreamDestroy(s->stream);
        delete s;
        return CTS_ERR_HIP;
    }
    float *sp = nullptr;
    if ((rc = scratch_ptr(&sp))) {
        cts_sur_destroy(s);
        return rc;
    }
    *out = s;
    return CTS_OK;
}

int cts_sur_destroy(cts_surrogate *s)
{
    if (!s) return CTS_OK;
    (void)hipStreamSynchronize(s->stream);
    void *p[] = {s->d_params, s->d_m, s->d_v, s->d_part, s->d_grad, s->d_loss, s->d_samples, s->d_val, s->d_err};
    for (void *q : p)
        if (q) (void)hipFree(q);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return CTS_OK;
}

int cts_sur_init_params(cts_surrogate *s, uint64_t seed)
{
    if (!s) return sfail(CTS_ERR_ARG, "null surrogate");
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t st = seed;
    auto uni = [&](double a) { return (float)(a * (2.0 * ((mix(st) >> 11) * 0x1p-53) - 1.0)); };
    for (int i = 0; i < 192; i++) p[MP_W1 + i] = uni(1.0 / std::sqrt(3.0));
    for (int i = 0; i < 4096; i++) p[MP_W2 + i] = uni(1.0 / 8.0);
    for (int i = 0; i < 128; i++) p[MP_W3 + i] = uni(1.0 / 8.0);
    SCHK(hipMemcpy(s->d_params, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    SCHK(hipMemset(s->d_m, 0, kMlpParams * sizeof(float)));
    SCHK(hipMemset(s->d_v, 0, kMlpParams * sizeof(float)));
    s->step = 0;
    return CTS_OK;
}

int cts_sur_set_params(cts_surrogate *s, const float *params, const float *norm)
{
    if (!s || !params || !norm) return sfail(CTS_ERR_ARG, "cts_sur_set_params: null");
    SCHK(hipMemcpy(s->d_params, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    for (int i = 0; i < 4; i++) s->norm[i] = norm[i];
    return CTS_OK;
}

int cts_sur_get_params(cts_surrogate *s, float *params, float *norm)
{
    if (!s) return sfail(CTS_ERR_ARG, "null surrogate");
    SCHK(hipStreamSynchronize(s->stream));
    if (params) SCHK(hipMemcpy(params, s->d_params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if (norm)
        for (int i = 0; i < 4; i++) norm[i] = s->norm[i];
    return CTS_OK;
}

int cts_sur_generate(cts_surrogate *s, int64_t n_train, int64_t n_val, uint64_t seed)
{
    if (!s || n_train < 64 || n_val < 1) return sfail(CTS_ERR_ARG, "cts_sur_generate: n_train >= 64, n_val >= 1");
    int rc;
    if ((rc = salloc(&s->d_samples, (size_t)n_train * 5)) || (rc = salloc(&s->d_val, (size_t)n_val * 5)) ||
        (rc = salloc(&s->d_err, (size_t)n_val)))
        return rc;
    s->n_train = n_train;
    s->n_val = n_val;
    SCHK(launch_mlp_samples(n_train, seed, 0x53414d50u, s->d_samples, nullptr, s->stream));
    float *d_tmp = nullptr;
    if ((rc = salloc(&d_tmp, (size_t)n_val * 5))) return rc;
    SCHK(launch_mlp_samples(n_val, seed, 0x56414c44u, d_tmp, s->d_val, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    (void)hipFree(d_tmp);
    // target normalisation from the training set (mean, standard deviation)
    std::vector<float> h((size_t)n_train * 5);
    SCHK(hipMemcpy(h.data(), s->d_samples, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int c = 0; c < 2; c++) {
        double sum = 0, sq = 0;
        for (int64_t i = 0; i < n_train; i++) sum += h[(size_t)i * 5 + 3 + c];
        const double mu = sum / n_train;
        for (int64_t i = 0; i < n_train; i++) {
            const double d = h[(size_t)i * 5 + 3 + c] - mu;
            sq += d * d;
        }
        const double sd = std::sqrt(sq / n_train);
        if (!std::isfinite(mu) || !(sd > 0)) return sfail(CTS_ERR_STATE, "cts_sur_generate: degenerate targets");
        s->norm[2 * c] = (float)mu;
        s->norm[2 * c + 1] = (float)sd;
    }
    return CTS_OK;
}

int cts_sur_get_samples(cts_surrogate *s, int64_t first, int64_t n, float *train5, double *val5)
{
    if (!s) return sfail(CTS_ERR_ARG, "null surrogate");
    SCHK(hipStreamSynchronize(s->stream));
    if (train5) {
        if (first < 0 || first + n > s->n_train) return sfail(CTS_ERR_ARG, "cts_sur_get_samples: range");
        SCHK(hipMemcpy(train5, s->d_samples + first * 5, (size_t)n * 5 * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (val5) {
        if (first < 0 || first + n > s->n_val) return sfail(CTS_ERR_ARG, "cts_sur_get_samples: range");
        SCHK(hipMemcpy(val5, s->d_val + first * 5, (size_t)n * 5 * sizeof(double), hipMemcpyDeviceToHost));
    }
    return CTS_OK;
}

int cts_sur_train(cts_surrogate *s, int32_t steps, int32_t nwg, int32_t chunks, float lr_max, float lr_min,
                  uint64_t seed, float *loss_out)
{
    if (!s || steps < 0 || nwg < 1 || chunks < 1 || !(lr_max > 0)) return sfail(CTS_ERR_ARG, "cts_sur_train");
    if (!s->d_samples) return sfail(CTS_ERR_STATE, "cts_sur_train: no samples (cts_sur_generate)");
    int rc = ensure_part(s, nwg);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float inv_batch = 1.0f / ((float)nwg * chunks * 64);
    for (int32_t t = 0; t < steps; t++) {
        // cosine schedule from lr_max to lr_min over this call
        const double c = steps > 1 ? 0.5 * (1.0 + std::cos(kPi * t / (steps - 1))) : 1.0;
        const float lr = (float)(lr_min + (lr_max - lr_min) * c);
        s->step++;
        const float bc1 = 1.0f - std::pow(b1, (float)s->step), bc2 = 1.0f - std::pow(b2, (float)s->step);
        SCHK(launch_mlp_train(s->d_params, s->d_samples, s->n_train, s->d_part, s->norm, inv_batch, seed,
                              (uint32_t)s->step, chunks, 1, nwg, s->stream));
        SCHK(launch_mlp_adam(s->d_params, s->d_m, s->d_v, s->d_part, nwg, lr, b1, b2, eps, bc1, bc2, nullptr,
                             s->d_loss, s->stream));
    }
    // returns with the parameters final (tracking streams read them)
    float l = 0;
    SCHK(hipMemcpyAsync(&l, s->d_loss, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (loss_out) *loss_out = l * inv_batch;
    return CTS_OK;
}

int cts_sur_gradient(cts_surrogate *s, const float *samples5, int64_t n, float *grad, float *loss)
{
    if (!s || !samples5 || n < 1 || !grad) return sfail(CTS_ERR_ARG, "cts_sur_gradient");
    const int chunks = 4;
    const int nwg = (int)((n + 64 * chunks - 1) / (64 * chunks));
    int rc = ensure_part(s, nwg);
    if (rc) return rc;
    float *d = nullptr;
    if ((rc = salloc(&d, (size_t)n * 5))) return rc;
    SCHK(hipMemcpy(d, samples5, (size_t)n * 5 * sizeof(float), hipMemcpyHostToDevice));
    const float inv_batch = 1.0f / (float)n;
    hipError_t e = launch_mlp_train(s->d_params, d, n, s->d_part, s->norm, inv_batch, 0, 0, chunks, 0, nwg, s->stream);
    if (e == hipSuccess)
        e = launch_mlp_adam(s->d_params, s->d_m, s->d_v, s->d_part, nwg, 0.0f, 0, 0, 0, 1, 1, s->d_grad, s->d_loss,
                            s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return sfail(CTS_ERR_HIP, std::string("gradient: ") + hipGetErrorString(e));
    SCHK(hipMemcpy(grad, s->d_grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) {
        float l = 0;
        SCHK(hipMemcpy(&l, s->d_loss, sizeof(float), hipMemcpyDeviceToHost));
        *loss = l * inv_batch;
    }
    return CTS_OK;
}

int cts_sur_eval(cts_surrogate *s, const float *features3, int64_t n, float *out2)
{
    if (!s || !features3 || !out2 || n < 1) return sfail(CTS_ERR_ARG, "cts_sur_eval");
    float *d = nullptr, *o = nullptr;
    int rc;
    if ((rc = salloc(&d, (size_t)n * 3)) || (rc = salloc(&o, (size_t)n * 2))) return rc;
    hipError_t e = hipMemcpy(d, features3, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_mlp_eval(s->d_params, d, n, 3, o, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(out2, o, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    (void)hipFree(o);
    if (e != hipSuccess) return sfail(CTS_ERR_HIP, std::string("eval: ") + hipGetErrorString(e));
    return CTS_OK;
}

int cts_sur_eval_noalloc(cts_surrogate *s, const float *features3, int64_t n, float *out2)
{
    if (!s || !features3 || !out2 || n < 1) return sfail(CTS_ERR_ARG, "cts_sur_eval_noalloc");
    if (n > kSurScratchPoints)
        return sfail(CTS_ERR_ARG, "cts_sur_eval_noalloc: n above " + std::to_string(kSurScratchPoints) +
                                      " points (the static staging buffer); call it in pieces or use cts_sur_eval");
    float *d = nullptr;
    int rc = scratch_ptr(&d);
    if (rc) return rc;
    // features in [0, 3n), outputs in [3n, 5n): every element used is
    // written by this call before it is read, nothing needs clearing
    float *o = d + 3 * n;
    SCHK(hipMemcpyAsync(d, features3, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SCHK(launch_mlp_eval(s->d_params, d, n, 3, o, s->stream));
    SCHK(hipMemcpyAsync(out2, o, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return CTS_OK;
}

int cts_sur_validate(cts_surrogate *s, double *max_err, double *rms_err, double *err_out)
{
    if (!s || !s->d_val) return sfail(CTS_ERR_STATE, "cts_sur_validate: no validation set");
    SCHK(launch_mlp_field_error(s->d_params, s->d_val, s->n_val, s->norm, s->d_err, s->stream));
    std::vector<double> e((size_t)s->n_val);
    SCHK(hipMemcpyAsync(e.data(), s->d_err, e.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    double mx = 0, sq = 0;
    for (double v : e) {
        mx = std::fmax(mx, v);
        sq += v * v;
    }
    if (max_err) *max_err = mx;
    if (rms_err) *rms_err = std::sqrt(sq / (double)e.size());
    if (err_out) std::memcpy(err_out, e.data(), e.size() * sizeof(double));
    return CTS_OK;
}

}  // extern "C"
