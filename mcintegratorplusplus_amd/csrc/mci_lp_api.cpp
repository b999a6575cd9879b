// mci_lp_api.cpp — C-ABI of the learned-proposal path (include/mci_amd.h, mci_mlp_* and
// mci_lp_*): the network (parameters, BF16 MFMA forward/backward, Adam) and the sampler that
// runs the plain random-walk chain (FP64 or FP32), collects training pairs from it, trains the
// network on the GPU and runs the chain that mixes random-walk and learned moves.
#include "mci_host.hpp"
#include "mci_em.hpp"
#include "mci_kernels.hpp"
#include "mci_lp.hpp"
#include "mci_mlp.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

using mcih::DBuf;
using mcih::det_sum;
using mcih::fail;
using mcim::bf16_t;

namespace
{
uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void philox_host(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t o[4])
{
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    o[0] = c0;
    o[1] = c1;
    o[2] = c2;
    o[3] = c3;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }
} // namespace

// ---- the network ----
struct mci_mlp
{
    int d = 0, H = 0, K = 0, O = 0, Op = 0;
    int64_t np = 0, npad = 0; // flat (unpadded) and padded parameter counts
    int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    DBuf<float> p, m, v, g;
    DBuf<bf16_t> pb;
    int64_t cap = 0;
    DBuf<bf16_t> Xb, H1b, H2b, dOb, dZ2b, dZ1b;
    DBuf<float> X, Y, Z1, Z2, OUT, dO, dH2, dH1, loss;
    DBuf<float> ws; // split-reduction partials of the weight gradients
    int adam_t = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ~mci_mlp()
    {
        if (own_stream && stream) { (void)hipStreamDestroy(stream); }
    }
};

namespace
{
#define CHECKM(x)                                                                                        \
    do {                                                                                                 \
        if (!(x)) { return fail(MCI_E_INVALID_ARGUMENT, "null network"); }                               \
    } while (0)

// flat (unpadded) <-> padded parameter layouts: only W3 and b3 differ (O -> Op columns)
void flat_to_padded(const mci_mlp *n, const float *f, std::vector<float> &pad)
{
    pad.assign((size_t)n->npad, 0.f);
    const int d = n->d, H = n->H, O = n->O, Op = n->Op;
    std::copy(f, f + (int64_t)d * H + H + (int64_t)H * H + H, pad.begin());
    const float *w3 = f + (int64_t)d * H + H + (int64_t)H * H + H;
    for (int i = 0; i < H; ++i) { std::copy(w3 + (int64_t)i * O, w3 + (int64_t)i * O + O, pad.begin() + n->oW3 + (int64_t)i * Op); }
    std::copy(w3 + (int64_t)H * O, w3 + (int64_t)H * O + O, pad.begin() + n->ob3);
}

void padded_to_flat(const mci_mlp *n, const std::vector<float> &pad, float *f)
{
    const int d = n->d, H = n->H, O = n->O, Op = n->Op;
    std::copy(pad.begin(), pad.begin() + (int64_t)d * H + H + (int64_t)H * H + H, f);
    float *w3 = f + (int64_t)d * H + H + (int64_t)H * H + H;
    for (int i = 0; i < H; ++i) { std::copy(pad.begin() + n->oW3 + (int64_t)i * Op, pad.begin() + n->oW3 + (int64_t)i * Op + O, w3 + (int64_t)i * O); }
    std::copy(pad.begin() + n->ob3, pad.begin() + n->ob3 + O, w3 + (int64_t)H * O);
}

mci_status mlp_upload_params(mci_mlp *n, const std::vector<float> &pad)
{
    HIPCHK(hipMemcpy(n->p.p, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<bf16_t> b(pad.size());
    for (size_t i = 0; i < pad.size(); ++i) {
        uint32_t u;
        std::memcpy(&u, &pad[i], 4);
        b[i] = (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    HIPCHK(hipMemcpy(n->pb.p, b.data(), b.size() * sizeof(bf16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(n->m.p, 0, pad.size() * sizeof(float)));
    HIPCHK(hipMemset(n->v.p, 0, pad.size() * sizeof(float)));
    n->adam_t = 0;
    return MCI_OK;
}

mci_status mlp_reserve(mci_mlp *n, int64_t cap)
{
    if (cap <= n->cap) { return MCI_OK; }
    const int d = n->d, H = n->H, Op = n->Op;
    HIPCHK(n->X.ensure((size_t)d * cap));
    HIPCHK(n->Y.ensure((size_t)d * cap));
    HIPCHK(n->Xb.ensure((size_t)d * cap));
    HIPCHK(n->Z1.ensure((size_t)H * cap));
    HIPCHK(n->H1b.ensure((size_t)H * cap));
    HIPCHK(n->Z2.ensure((size_t)H * cap));
    HIPCHK(n->H2b.ensure((size_t)H * cap));
    HIPCHK(n->OUT.ensure((size_t)Op * cap));
    HIPCHK(n->dO.ensure((size_t)Op * cap));
    HIPCHK(n->dOb.ensure((size_t)Op * cap));
    HIPCHK(n->dH2.ensure((size_t)H * cap));
    HIPCHK(n->dZ2b.ensure((size_t)H * cap));
    HIPCHK(n->dH1.ensure((size_t)H * cap));
    HIPCHK(n->dZ1b.ensure((size_t)H * cap));
    HIPCHK(n->loss.ensure((size_t)cap));
    HIPCHK(n->ws.ensure((size_t)32 * std::max({(size_t)H * Op, (size_t)H * H, (size_t)d * H})));
    n->cap = cap;
    return MCI_OK;
}

// forward of n samples from BF16 inputs xb [d][n] into out [Op][n] (feature-major)
mci_status mlp_forward(mci_mlp *n, int64_t ns, const bf16_t *xb, float *out)
{
    const int d = n->d, H = n->H, Op = n->Op;
    const bf16_t *W = n->pb.p;
    const float *P = n->p.p;
    hipStream_t s = n->stream;
    HIPCHK(mcim::gemm_bf16(true, false, mcim::EPI_BIAS_SILU, H, (int)ns, d, W + n->oW1, xb, P + n->ob1, n->Z1.p, (int)ns, n->H1b.p,
                           (int)ns, s));
    HIPCHK(mcim::gemm_bf16(true, false, mcim::EPI_BIAS_SILU, H, (int)ns, H, W + n->oW2, n->H1b.p, P + n->ob2, n->Z2.p, (int)ns,
                           n->H2b.p, (int)ns, s));
    HIPCHK(mcim::gemm_bf16(true, false, mcim::EPI_BIAS, Op, (int)ns, H, W + n->oW3, n->H2b.p, P + n->ob3, out, (int)ns, nullptr, 0, s));
    return MCI_OK;
}

// one forward + backward on (X, Y) [d][ns] already on the device; gradient of the mean loss in n->g
mci_status mlp_fwd_bwd(mci_mlp *n, int64_t ns)
{
    const int d = n->d, H = n->H, K = n->K, Op = n->Op;
    hipStream_t s = n->stream;
    const int N = (int)ns;
    HIPCHK(mcim::launch_to_bf16_scaled(n->X.p, n->Xb.p, (int64_t)d * ns, mcim::IN_SCALE, s));
    TRY(mlp_forward(n, ns, n->Xb.p, n->OUT.p));
    HIPCHK(mcim::launch_mixture_head(n->OUT.p, N, n->Y.p, N, d, K, Op, n->loss.p, n->dO.p, n->dOb.p, 1.0f / (float)ns, s));
    float *g = n->g.p;
    const bf16_t *W = n->pb.p;
    HIPCHK(mcim::launch_rowsum(n->dO.p, Op, N, N, g + n->ob3, s));
    HIPCHK(mcim::gemm_bf16_splitk(false, true, H, Op, N, n->H2b.p, n->dOb.p, g + n->oW3, Op, n->ws.p, n->ws.n, s));
    HIPCHK(mcim::gemm_bf16(false, false, mcim::EPI_NONE, H, N, Op, W + n->oW3, n->dOb.p, nullptr, n->dH2.p, N, nullptr, 0, s));
    HIPCHK(mcim::launch_silu_back(n->dH2.p, n->Z2.p, n->dZ2b.p, (int64_t)H * ns, s));
    HIPCHK(mcim::launch_rowsum(n->dH2.p, H, N, N, g + n->ob2, s));
    HIPCHK(mcim::gemm_bf16_splitk(false, true, H, H, N, n->H1b.p, n->dZ2b.p, g + n->oW2, H, n->ws.p, n->ws.n, s));
    HIPCHK(mcim::gemm_bf16(false, false, mcim::EPI_NONE, H, N, H, W + n->oW2, n->dZ2b.p, nullptr, n->dH1.p, N, nullptr, 0, s));
    HIPCHK(mcim::launch_silu_back(n->dH1.p, n->Z1.p, n->dZ1b.p, (int64_t)H * ns, s));
    HIPCHK(mcim::launch_rowsum(n->dH1.p, H, N, N, g + n->ob1, s));
    HIPCHK(mcim::gemm_bf16_splitk(false, true, d, H, N, n->Xb.p, n->dZ1b.p, g + n->oW1, H, n->ws.p, n->ws.n, s));
    return MCI_OK;
}

mci_status mlp_adam(mci_mlp *n, float lr)
{
    ++n->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, (float)n->adam_t), bc2 = 1.0f - std::pow(b2, (float)n->adam_t);
    HIPCHK(mcim::launch_adam(n->p.p, n->m.p, n->v.p, n->g.p, n->pb.p, n->npad, lr, b1, b2, 1e-8f, bc1, bc2, n->stream));
    return MCI_OK;
}

mci_status mlp_make(int d, int H, int K, uint64_t seed, hipStream_t stream, mci_mlp **out)
{
    if (d < 1 || d > 256 || H < 16 || H > 1024 || K < 1 || K > 8) { return fail(MCI_E_INVALID_ARGUMENT, "network sizes out of range"); }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { return fail(MCI_E_DEVICE, "no HIP device available"); }
    mci_mlp *n = new mci_mlp();
    n->d = d;
    n->H = H;
    n->K = K;
    n->O = K * (1 + 2 * d);
    n->Op = round_up(n->O, 64);
    n->oW1 = 0;
    n->ob1 = n->oW1 + (int64_t)d * H;
    n->oW2 = n->ob1 + H;
    n->ob2 = n->oW2 + (int64_t)H * H;
    n->oW3 = n->ob2 + H;
    n->ob3 = n->oW3 + (int64_t)H * n->Op;
    n->npad = n->ob3 + n->Op;
    n->np = (int64_t)d * H + H + (int64_t)H * H + H + (int64_t)H * n->O + n->O;
    if (stream) { n->stream = stream; }
    else {
        if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
            delete n;
            return fail(MCI_E_DEVICE, "stream creation failed");
        }
        n->own_stream = true;
    }
    if (n->p.ensure(n->npad) != hipSuccess || n->m.ensure(n->npad) != hipSuccess || n->v.ensure(n->npad) != hipSuccess ||
        n->g.ensure(n->npad) != hipSuccess || n->pb.ensure(n->npad) != hipSuccess) {
        delete n;
        return fail(MCI_E_DEVICE, "allocation failed");
    }
    // Glorot-uniform weights (output layer scaled by 0.1), zero biases
    std::vector<float> flat((size_t)n->np, 0.f);
    uint64_t st = seed ^ 0x5DEECE66DULL;
    auto fill = [&](int64_t off, int fin, int fout, float gain) {
        const double a = gain * std::sqrt(6.0 / (fin + fout));
        for (int64_t i = 0; i < (int64_t)fin * fout; ++i) {
            const double u = (double)(splitmix(st) >> 11) * 0x1p-53;
            flat[(size_t)(off + i)] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(0, d, H, 1.f);
    fill((int64_t)d * H + H, H, H, 1.f);
    fill((int64_t)d * H + H + (int64_t)H * H + H, H, n->O, 0.1f);
    std::vector<float> pad;
    flat_to_padded(n, flat.data(), pad);
    const mci_status s = mlp_upload_params(n, pad);
    if (s != MCI_OK) {
        delete n;
        return s;
    }
    *out = n;
    return MCI_OK;
}

// host [n][d] -> device [d][n]
mci_status upload_transposed(float *dst, const float *src, int64_t n, int d)
{
    std::vector<float> t((size_t)n * d);
    for (int64_t r = 0; r < n; ++r) {
        for (int j = 0; j < d; ++j) { t[(size_t)j * n + r] = src[(size_t)r * d + j]; }
    }
    HIPCHK(hipMemcpy(dst, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    return MCI_OK;
}
} // namespace

// ---- the sampler ----
struct mci_lp
{
    int d = 0, K = 0;
    int64_t W = 0;
    uint64_t seed = 1;
    int H = 128, Kq = 8;
    std::vector<double> mu, P, U, Um, v, c, L; // padded to KMAX components
    std::vector<double> cw;
    bool dev = false;
    hipStream_t stream = nullptr;
    // target on the device
    DBuf<double> mu64, P64, U64, Um64, v64, c64;
    DBuf<float> mu32, P32, U32, Um32, v32, c32, Lf, muf, cwf, Ut32;
    // chain state (one precision at a time)
    DBuf<double> x64, q64, m64, S64, sb64, cb64;
    DBuf<float> x32, q32, m32, S32, sb32, cb32;
    DBuf<double> sq, bmeans, wavg, werr;
    DBuf<int> jb, bi;
    DBuf<int32_t> nacc, nacc_ind;
    // learned move
    mci_mlp *net = nullptr;
    DBuf<float> snap, nb1, nb2, nb3;
    DBuf<uint16_t> nW1t, nW2t, nW3t;
    uint64_t train_iter = 0; // minibatches drawn so far (each draws its own pairs)
    bool fixed_input = false; // train with the fixed input ctx0 (an independence proposal)
    DBuf<float> dctx0;        // ctx0 on the device
    DBuf<float> imix;         // the independence proposal's mixture (mu, 1 / sigma, sigma, lc, cumulative weights)
    int nsnap = 0;
    double same_frac = 0.5;
    std::vector<float> ctx0;   // mean of the last snapshot: the network input of the independence proposal
    uint64_t t_next = 0;
    double step64 = 0., step32 = 0.;
    double results[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double timings[4] = {0, 0, 0, 0};
    double loss_first = 0., loss_last = 0.;
    std::vector<double> h_wavg;
    ~mci_lp()
    {
        delete net;
        if (stream) { (void)hipStreamDestroy(stream); }
    }
};

namespace
{
#define CHECKL(x)                                                                                        \
    do {                                                                                                 \
        if (!(x)) { return fail(MCI_E_INVALID_ARGUMENT, "null sampler"); }                               \
    } while (0)

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

template <class T>
mci_status up(DBuf<T> &dst, const std::vector<T> &src)
{
    HIPCHK(dst.ensure(src.size()));
    HIPCHK(hipMemcpy(dst.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return MCI_OK;
}

template <class T>
std::vector<float> to_f(const std::vector<T> &a)
{
    return std::vector<float>(a.begin(), a.end());
}

mci_status lp_device(mci_lp *lp)
{
    if (lp->dev) { return MCI_OK; }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) { return fail(MCI_E_DEVICE, "no HIP device available"); }
    HIPCHK(hipStreamCreateWithFlags(&lp->stream, hipStreamNonBlocking));
    TRY(up(lp->mu64, lp->mu));
    TRY(up(lp->P64, lp->P));
    TRY(up(lp->U64, lp->U));
    TRY(up(lp->U32, to_f(lp->U)));
    {
        const int d = lp->d;
        std::vector<float> ut((size_t)mcil::KMAX * d * 32, 0.f);
        for (int k = 0; k < mcil::KMAX; ++k) {
            for (int j = 0; j < d; ++j) {
                for (int r = 0; r < d; ++r) { ut[((size_t)k * d + j) * 32 + r] = (float)lp->U[((size_t)k * d + r) * d + j]; }
            }
        }
        TRY(up(lp->Ut32, ut));
    }
    TRY(up(lp->Um64, lp->Um));
    TRY(up(lp->Um32, to_f(lp->Um)));
    TRY(up(lp->v64, lp->v));
    TRY(up(lp->c64, lp->c));
    std::vector<float> cf(lp->c.begin(), lp->c.end());
    TRY(up(lp->mu32, to_f(lp->mu)));
    TRY(up(lp->P32, to_f(lp->P)));
    TRY(up(lp->v32, to_f(lp->v)));
    TRY(up(lp->c32, cf));
    TRY(up(lp->Lf, to_f(lp->L)));
    TRY(up(lp->muf, to_f(lp->mu)));
    TRY(up(lp->cwf, to_f(lp->cw)));
    const size_t W = (size_t)lp->W;
    HIPCHK(lp->sq.ensure(W));
    HIPCHK(lp->jb.ensure(W));
    HIPCHK(lp->bi.ensure(W));
    HIPCHK(lp->nacc.ensure(W));
    HIPCHK(lp->nacc_ind.ensure(W));
    lp->dev = true;
    return MCI_OK;
}

template <class R>
mcil::Target<R> target_of(mci_lp *lp);
template <>
mcil::Target<double> target_of<double>(mci_lp *lp)
{
    return {lp->d, lp->mu64.p, lp->P64.p, lp->U64.p, lp->Um64.p, lp->v64.p, lp->c64.p, lp->Lf.p, lp->muf.p, lp->cwf.p, lp->K, lp->Ut32.p};
}
template <>
mcil::Target<float> target_of<float>(mci_lp *lp)
{
    return {lp->d, lp->mu32.p, lp->P32.p, lp->U32.p, lp->Um32.p, lp->v32.p, lp->c32.p, lp->Lf.p, lp->muf.p, lp->cwf.p, lp->K, lp->Ut32.p};
}

template <class R>
mci_status state_of(mci_lp *lp, int B, int nblocks, mcil::State<R> &S);
template <>
mci_status state_of<double>(mci_lp *lp, int B, int nblocks, mcil::State<double> &S)
{
    const size_t W = (size_t)lp->W;
    HIPCHK(lp->x64.ensure(W * lp->d));
    HIPCHK(lp->q64.ensure(W * mcil::KMAX));
    HIPCHK(lp->m64.ensure(W));
    HIPCHK(lp->S64.ensure(W));
    HIPCHK(lp->sb64.ensure(2 * W));
    HIPCHK(lp->cb64.ensure(2 * W));
    HIPCHK(lp->bmeans.ensure((size_t)std::max(nblocks, 1) * 2 * W));
    S = {lp->W, lp->x64.p, lp->q64.p, lp->m64.p, lp->S64.p, lp->sq.p, lp->sb64.p, lp->cb64.p, lp->jb.p, lp->bi.p, lp->bmeans.p, B, lp->nacc.p};
    return MCI_OK;
}
template <>
mci_status state_of<float>(mci_lp *lp, int B, int nblocks, mcil::State<float> &S)
{
    const size_t W = (size_t)lp->W;
    HIPCHK(lp->x32.ensure(W * lp->d));
    HIPCHK(lp->q32.ensure(W * mcil::KMAX));
    HIPCHK(lp->m32.ensure(W));
    HIPCHK(lp->S32.ensure(W));
    HIPCHK(lp->sb32.ensure(2 * W));
    HIPCHK(lp->cb32.ensure(2 * W));
    HIPCHK(lp->bmeans.ensure((size_t)std::max(nblocks, 1) * 2 * W));
    S = {lp->W, lp->x32.p, lp->q32.p, lp->m32.p, lp->S32.p, lp->sq.p, lp->sb32.p, lp->cb32.p, lp->jb.p, lp->bi.p, lp->bmeans.p, B, lp->nacc.p};
    return MCI_OK;
}

// MCI::findMRT2Step (src/MCIntegrator.cpp:87-168) with all walkers as ranks: MIN_STAT 100 steps
// per iteration (max(100, sqrt(40000 d)/W)), rate averaged over walkers, MIN_CONS 5, TOL 0.05,
// factor in [0.5, 2]
template <class R>
mci_status calibrate(mci_lp *lp, mcil::State<R> S, double &step)
{
    const mcil::Target<R> T = target_of<R>(lp);
    S.B = 0;
    const int64_t min_stat = (int64_t)std::fmax(100., std::sqrt(40000. * lp->d) / (double)lp->W);
    std::vector<int32_t> acc((size_t)lp->W);
    std::vector<double> rates((size_t)lp->W);
    int cons = 0, counter = 0;
    const uint32_t k0 = (uint32_t)lp->seed, k1 = (uint32_t)(lp->seed >> 32);
    while (cons < 5) {
        HIPCHK(hipMemsetAsync(lp->nacc.p, 0, sizeof(int32_t) * lp->W, lp->stream));
        HIPCHK(mcil::launch_lp_rw<R>(T, S, (R)step, lp->t_next, min_stat, k0, k1, lp->stream));
        lp->t_next += (uint64_t)min_stat;
        HIPCHK(hipMemcpyAsync(acc.data(), lp->nacc.p, sizeof(int32_t) * lp->W, hipMemcpyDeviceToHost, lp->stream));
        HIPCHK(hipStreamSynchronize(lp->stream));
        for (int64_t w = 0; w < lp->W; ++w) {
            rates[(size_t)w] = acc[(size_t)w] > 0 ? (double)acc[(size_t)w] / (double)min_stat : 0.;
        }
        const double rate = det_sum(rates.data(), lp->W) / (double)lp->W;
        if (std::fabs(rate - 0.5) < 0.05) { ++cons; }
        else { cons = 0; }
        step = step * std::fmin(2., std::fmax(0.5, rate / 0.5));
        step = std::fmin(std::fmax(step, (double)FLT_MIN), (double)FLT_MAX);
        if (++counter >= 50) { break; }
    }
    return MCI_OK;
}

mci_status net_ready(mci_lp *lp)
{
    if (!lp->net) { TRY(mlp_make(lp->d, lp->H, lp->Kq, lp->seed, lp->stream, &lp->net)); }
    return MCI_OK;
}

// per-walker MJBlocker over the block means, then avg, the standard error of the walker means
// and the MPIMCI combine (src/MPIMCI.cpp:85-92), both observables
mci_status estimate(mci_lp *lp, int nblocks)
{
    const size_t W = (size_t)lp->W;
    HIPCHK(lp->wavg.ensure(2 * W));
    HIPCHK(lp->werr.ensure(2 * W));
    mcid::EstimJob j;
    j.data = lp->bmeans.p;
    j.scratch = nullptr;
    j.n = nblocks;
    j.W = lp->W;
    j.nobs = 2;
    j.off = 0;
    j.estim = MCI_ESTIM_MJBLOCKER;
    j.scale = 1.0;
    j.active = nullptr;
    j.avg = lp->wavg.p;
    j.err = lp->werr.p;
    HIPCHK(mcid::launch_estimate(j, lp->stream));
    std::vector<double> a(2 * W), e(2 * W), tmp(W);
    HIPCHK(hipMemcpyAsync(a.data(), lp->wavg.p, sizeof(double) * 2 * W, hipMemcpyDeviceToHost, lp->stream));
    HIPCHK(hipMemcpyAsync(e.data(), lp->werr.p, sizeof(double) * 2 * W, hipMemcpyDeviceToHost, lp->stream));
    HIPCHK(hipStreamSynchronize(lp->stream));
    for (int o = 0; o < 2; ++o) {
        const double *ao = a.data() + o * W, *eo = e.data() + o * W;
        const double avg = det_sum(ao, (int64_t)W) / (double)W;
        for (size_t w = 0; w < W; ++w) { tmp[w] = (ao[w] - avg) * (ao[w] - avg); }
        const double spread = std::sqrt(det_sum(tmp.data(), (int64_t)W) / ((double)W * (double)(W - 1)));
        for (size_t w = 0; w < W; ++w) { tmp[w] = eo[w] * eo[w]; }
        const double comb = std::sqrt(det_sum(tmp.data(), (int64_t)W)) / (double)W;
        lp->results[o] = avg;
        lp->results[2 + o] = spread;
        lp->results[4 + o] = comb;
    }
    lp->h_wavg = a;
    return MCI_OK;
}

template <class R>
mci_status run_plain(mci_lp *lp, int64_t nsteps, int nblocks)
{
    mcil::State<R> S;
    const int B = (int)(nsteps / nblocks);
    TRY(state_of<R>(lp, B, nblocks, S));
    const mcil::Target<R> T = target_of<R>(lp);
    const uint32_t k0 = (uint32_t)lp->seed, k1 = (uint32_t)(lp->seed >> 32);
    auto t0 = clk::now();
    HIPCHK(mcil::launch_lp_init<R>(T, S, k0, k1, lp->stream));
    double &step = (sizeof(R) == 8) ? lp->step64 : lp->step32;
    if (step <= 0.) { step = 1.0; }
    TRY(calibrate<R>(lp, S, step));
    lp->timings[0] = ms_since(t0);
    HIPCHK(hipMemsetAsync(lp->nacc.p, 0, sizeof(int32_t) * lp->W, lp->stream));
    HIPCHK(hipStreamSynchronize(lp->stream));
    t0 = clk::now();
    HIPCHK(mcil::launch_lp_rw<R>(T, S, (R)step, lp->t_next, nsteps, k0, k1, lp->stream));
    lp->t_next += (uint64_t)nsteps;
    HIPCHK(hipStreamSynchronize(lp->stream));
    lp->timings[1] = ms_since(t0);
    t0 = clk::now();
    TRY(estimate(lp, nblocks));
    lp->timings[2] = ms_since(t0);
    std::vector<int32_t> acc((size_t)lp->W);
    HIPCHK(hipMemcpy(acc.data(), lp->nacc.p, sizeof(int32_t) * lp->W, hipMemcpyDeviceToHost));
    double a = 0.;
    for (int32_t v : acc) { a += v; }
    lp->results[6] = a / ((double)lp->W * (double)nsteps);
    lp->results[7] = 0.;
    return MCI_OK;
}

// the network in the layout of the fused learned move (BF16 [out][in] weights)
mci_status net_dev(mci_lp *lp, mcil::NetDev &N)
{
    TRY(net_ready(lp));
    mci_mlp *n = lp->net;
    const int d = n->d, H = n->H, K = n->K, dp = std::max(16, d);
    if (!mcil::learned_supported(d, H, K)) {
        return fail(MCI_E_INVALID_ARGUMENT, "the learned move needs d in {8, 16, 32}, hidden in {64, 128} and 1 <= ncomp <= %d", mcil::KMAX);
    }
    const int Op = 2 * d * mcil::KMAX + 32, L0 = 2 * d * K;
    std::vector<float> p((size_t)n->npad);
    HIPCHK(hipStreamSynchronize(n->stream));
    HIPCHK(hipMemcpy(p.data(), n->p.p, p.size() * sizeof(float), hipMemcpyDeviceToHost));
    auto bf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    };
    std::vector<uint16_t> w1((size_t)H * dp, 0), w2((size_t)H * H), w3((size_t)Op * H, 0);
    std::vector<float> b3((size_t)Op, 0.f);
    for (int h = 0; h < H; ++h) {
        for (int i = 0; i < d; ++i) { w1[(size_t)h * dp + i] = bf(p[(size_t)(n->oW1 + (int64_t)i * H + h)]); }
        for (int i = 0; i < H; ++i) { w2[(size_t)h * H + i] = bf(p[(size_t)(n->oW2 + (int64_t)i * H + h)]); }
    }
    // network output row o -> kernel row: (component k, mean j) -> 2 d k + 2 j, (k, log-scale j)
    // -> 2 d k + 2 j + 1, logit k -> 2 d KMAX + k
    auto dst = [&](int o) {
        if (o >= L0) { return 2 * d * mcil::KMAX + (o - L0); }
        const int k = o / (2 * d), r = o % (2 * d);
        return 2 * d * k + (r < d ? 2 * r : 2 * (r - d) + 1);
    };
    for (int o = 0; o < n->O; ++o) {
        const int r = dst(o);
        for (int h = 0; h < H; ++h) { w3[(size_t)r * H + h] = bf(p[(size_t)(n->oW3 + (int64_t)h * n->Op + o)]); }
        b3[(size_t)r] = p[(size_t)(n->ob3 + o)];
    }
    TRY(up(lp->nW1t, w1));
    TRY(up(lp->nW2t, w2));
    TRY(up(lp->nW3t, w3));
    TRY(up(lp->nb1, std::vector<float>(p.begin() + n->ob1, p.begin() + n->ob1 + H)));
    TRY(up(lp->nb2, std::vector<float>(p.begin() + n->ob2, p.begin() + n->ob2 + H)));
    TRY(up(lp->nb3, b3));
    N = {d, dp, H, K, n->O, Op, lp->nW1t.p, lp->nW2t.p, lp->nW3t.p, lp->nb1.p, lp->nb2.p, lp->nb3.p};
    return MCI_OK;
}

// the independence proposal: the network's mixture at ctx0 (the mean of the last snapshot), as the
// fused kernel evaluates it (means and 1 / sigma rounded to halves), into lp->imix: mu, 1 / sigma,
// sigma [KMAX][d], lc [KMAX] (log weight + sum log(1 / sigma) - d log(2 pi) / 2, -inf past K),
// cumulative weights [KMAX]; with wts, also the normalised weights [K] in FP64
mci_status indep_mixture(mci_lp *lp, const mcil::NetDev &N, mcil::IndepMix &M, std::vector<double> *wts = nullptr)
{
    if (lp->ctx0.empty()) { return fail(MCI_E_RUNTIME, "no training samples: call mci_lp_collect and mci_lp_train first"); }
    const int d = lp->d, K = N.K, KM = mcil::KMAX;
    HIPCHK(lp->dctx0.ensure((size_t)d));
    HIPCHK(hipMemcpy(lp->dctx0.p, lp->ctx0.data(), sizeof(float) * (size_t)d, hipMemcpyHostToDevice));
    DBuf<float> dout;
    HIPCHK(dout.ensure((size_t)N.O));
    mcil::State<float> S1{};
    S1.W = 1;
    HIPCHK(mcil::launch_lp_network(N, target_of<float>(lp), S1, lp->dctx0.p, dout.p, lp->stream));
    std::vector<float> o((size_t)N.O);
    HIPCHK(hipStreamSynchronize(lp->stream));
    HIPCHK(hipMemcpy(o.data(), dout.p, sizeof(float) * o.size(), hipMemcpyDeviceToHost));
    const size_t KD = (size_t)KM * d;
    std::vector<float> h(3 * KD + 2 * KM, 0.f);
    float *mu = h.data(), *isig = mu + KD, *sig = isig + KD, *lc = sig + KD, *cw = lc + KM;
    double amax = -HUGE_VAL, sa = 0., cum = 0.;
    for (int k = 0; k < K; ++k) { amax = std::fmax(amax, (double)o[(size_t)2 * d * K + k]); }
    for (int k = 0; k < K; ++k) { sa += std::exp((double)o[(size_t)2 * d * K + k] - amax); }
    for (int k = 0; k < KM; ++k) {
        if (k >= K) {
            for (int j = 0; j < d; ++j) { isig[(size_t)k * d + j] = sig[(size_t)k * d + j] = 1.f; }
            lc[k] = -HUGE_VALF;
            cw[k] = 1.f;
            continue;
        }
        const double lw = (double)o[(size_t)2 * d * K + k] - amax - std::log(sa);
        double ls = 0.;
        for (int j = 0; j < d; ++j) {
            mu[(size_t)k * d + j] = o[(size_t)2 * d * k + j];
            const float is = (float)std::exp(-(double)o[(size_t)2 * d * k + d + j]);
            isig[(size_t)k * d + j] = is;
            sig[(size_t)k * d + j] = (float)(1.0 / (double)is);
            ls += std::log((double)is);
        }
        lc[k] = (float)(lw + ls - 0.5 * d * std::log(2.0 * M_PI));
        cum += std::exp(lw);
        if (wts) { wts->push_back(std::exp(lw)); }
        cw[k] = k == K - 1 ? 1.f : (float)cum;
    }
    TRY(up(lp->imix, h));
    const float *b = lp->imix.p;
    M = {b, b + KD, b + 2 * KD, b + 3 * KD, b + 3 * KD + KM, K};
    return MCI_OK;
}

// the chain mixing random-walk moves with learned moves: cycles of LP_CYCLE steps, the first
// LP_CYCLE - nl random-walk moves (one launch), then nl = round(pmix LP_CYCLE) learned moves (one
// fused launch).  Each move leaves p invariant, so does the cycle.  The learned moves use the
// conditional proposal q(x' | x), or with indep the independence proposal q(x'): the network's
// mixture at ctx0, the mean of the last snapshot.
constexpr int LP_CYCLE = 32;
mci_status run_mixed(mci_lp *lp, int64_t nsteps, int nblocks, double pmix, bool indep)
{
    mcil::NetDev N;
    TRY(net_dev(lp, N));
    mcil::IndepMix M{};
    if (indep) { TRY(indep_mixture(lp, N, M)); }
    mcil::State<float> S;
    const int B = (int)(nsteps / nblocks);
    TRY(state_of<float>(lp, B, nblocks, S));
    const mcil::Target<float> T = target_of<float>(lp);
    const uint32_t k0 = (uint32_t)lp->seed, k1 = (uint32_t)(lp->seed >> 32);
    const size_t W = (size_t)lp->W;
    auto t0 = clk::now();
    HIPCHK(mcil::launch_lp_init<float>(T, S, k0, k1, lp->stream));
    if (lp->step32 <= 0.) { lp->step32 = 1.0; }
    TRY(calibrate<float>(lp, S, lp->step32));
    lp->timings[0] = ms_since(t0);
    HIPCHK(hipMemsetAsync(lp->nacc.p, 0, sizeof(int32_t) * W, lp->stream));
    HIPCHK(hipMemsetAsync(lp->nacc_ind.p, 0, sizeof(int32_t) * W, lp->stream));
    HIPCHK(hipStreamSynchronize(lp->stream));
    t0 = clk::now();
    const uint64_t tb = lp->t_next;
    // cycles of LP_CYCLE steps: random-walk moves first, then learned moves (one launch each)
    const int nl = std::min(LP_CYCLE, std::max(1, (int)std::lround(pmix * LP_CYCLE))), nr = LP_CYCLE - nl;
    int64_t nind = 0, s = 0;
    while (s < nsteps) {
        const int64_t r = std::min<int64_t>(nr, nsteps - s);
        if (r > 0) {
            HIPCHK(mcil::launch_lp_rw<float>(T, S, (float)lp->step32, tb + (uint64_t)s, r, k0, k1, lp->stream));
            s += r;
        }
        const int64_t l = std::min<int64_t>(nl, nsteps - s);
        if (l > 0) {
            if (indep) { HIPCHK(mcil::launch_lp_indep(M, T, S, tb + (uint64_t)s, (int)l, k0, k1, lp->nacc_ind.p, lp->stream)); }
            else { HIPCHK(mcil::launch_lp_learned(N, T, S, tb + (uint64_t)s, (int)l, k0, k1, lp->nacc_ind.p, lp->stream)); }
            nind += l;
            s += l;
        }
    }
    lp->t_next += (uint64_t)nsteps;
    HIPCHK(hipStreamSynchronize(lp->stream));
    lp->timings[1] = ms_since(t0);
    t0 = clk::now();
    TRY(estimate(lp, nblocks));
    lp->timings[2] = ms_since(t0);
    std::vector<int32_t> acc(W), acci(W);
    HIPCHK(hipMemcpy(acc.data(), lp->nacc.p, sizeof(int32_t) * W, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(acci.data(), lp->nacc_ind.p, sizeof(int32_t) * W, hipMemcpyDeviceToHost));
    double a = 0., ai = 0.;
    for (size_t w = 0; w < W; ++w) {
        a += acc[w];
        ai += acci[w];
    }
    const int64_t nrw = nsteps - nind;
    lp->results[6] = nrw > 0 ? a / ((double)lp->W * (double)nrw) : 0.;
    lp->results[7] = nind > 0 ? ai / ((double)lp->W * (double)nind) : 0.;
    return MCI_OK;
}

// mci_gmm_fit_diag's fit with its sample sums on the device (stream s, current device): the same
// greedy k-means++ seeding -- the host draws the candidates from the same engine, the device
// computes the squared distances (bit-identical to the host fit's) and the candidates' potentials
// -- then the EM iterations on the device (mci_em.hip).  x: host [n][d].
mci_status gmm_fit_device(const double *x, int64_t n, int d, int K, int em_iters, uint64_t seed, hipStream_t s, double *mu_out,
                          double *var_out, double *pi_out, double *loglik)
{
    using namespace mcie;
    if (!x || !mu_out || !var_out || !pi_out || n < 1 || d < 1 || d > EM_DMAX || K < 1 || K > EM_KMAX || n < K) {
        return fail(MCI_E_INVALID_ARGUMENT, "need x, outputs, 1 <= d <= %d, 1 <= K <= %d and n >= K", EM_DMAX, EM_KMAX);
    }
    if (em_iters < 0 || em_iters > 1000) { return fail(MCI_E_INVALID_ARGUMENT, "EM iterations must be in [0, 1000]"); }
    const int64_t nb = em_nblocks(n);
    const size_t ne = (size_t)K * (1 + 2 * d);
    DBuf<double> dx, d2, part, dpi, dmu, dvar, dlc;
    DBuf<int64_t> dcand;
    HIPCHK(dx.ensure((size_t)n * d));
    HIPCHK(d2.ensure((size_t)n));
    HIPCHK(part.ensure((size_t)nb * std::max<size_t>(EM_NCAND, ne)));
    HIPCHK(dpi.ensure((size_t)K));
    HIPCHK(dmu.ensure((size_t)K * d));
    HIPCHK(dvar.ensure((size_t)K * d));
    HIPCHK(dlc.ensure((size_t)K));
    HIPCHK(dcand.ensure(EM_NCAND));
    HIPCHK(hipMemcpyAsync(dx.p, x, sizeof(double) * (size_t)n * d, hipMemcpyHostToDevice, s));
    // the initial width: the samples' mean squared deviation (the host fit's loops)
    std::vector<double> mean(d, 0.);
    double tot = 0.;
    for (int64_t r = 0; r < n; ++r) {
        for (int j = 0; j < d; ++j) { mean[j] += x[(size_t)r * d + j] / (double)n; }
    }
    for (int64_t r = 0; r < n; ++r) {
        for (int j = 0; j < d; ++j) { const double t = x[(size_t)r * d + j] - mean[j]; tot += t * t; }
    }
    const double v0 = std::max(tot / ((double)n * d), 1e-6);
    std::mt19937_64 g(seed ^ 0x5DEECE66DULL);
    std::uniform_real_distribution<double> U(0., 1.);
    std::vector<double> mu((size_t)K * d), var((size_t)K * d, v0), pi((size_t)K, 1.0 / K), d2h((size_t)n), ph((size_t)nb * EM_NCAND);
    int64_t pick = (int64_t)(U(g) * (double)n) % n;
    for (int k = 0; k < K; ++k) {
        for (int j = 0; j < d; ++j) { mu[(size_t)k * d + j] = x[(size_t)pick * d + j]; }
        HIPCHK(launch_em_dist(dx.p, n, d, pick, d2.p, k == 0, s));
        if (k == K - 1) { break; }
        HIPCHK(hipMemcpyAsync(d2h.data(), d2.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double sum = 0.;
        for (int64_t r = 0; r < n; ++r) { sum += d2h[(size_t)r]; }
        // the candidates: the first sample whose running sum of d2 exceeds u sum, for EM_NCAND draws
        // (one pass over the samples with the draws in increasing order)
        double u[EM_NCAND];
        int64_t cand[EM_NCAND];
        int ord[EM_NCAND];
        for (int l = 0; l < EM_NCAND; ++l) {
            u[l] = U(g) * sum;
            cand[l] = n - 1;
            ord[l] = l;
        }
        std::sort(ord, ord + EM_NCAND, [&](int a, int b) { return u[a] < u[b]; });
        double c = 0.;
        int nx = 0;
        for (int64_t r = 0; r < n && nx < EM_NCAND; ++r) {
            c += d2h[(size_t)r];
            while (nx < EM_NCAND && u[ord[nx]] < c) { cand[ord[nx++]] = r; }
        }
        HIPCHK(hipMemcpyAsync(dcand.p, cand, sizeof(cand), hipMemcpyHostToDevice, s));
        HIPCHK(launch_em_potential(dx.p, n, d, d2.p, dcand.p, EM_NCAND, part.p, s));
        HIPCHK(hipMemcpyAsync(ph.data(), part.p, sizeof(double) * ph.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double best = HUGE_VAL;
        for (int l = 0; l < EM_NCAND; ++l) {
            double pot = 0.;
            for (int64_t b = 0; b < nb; ++b) { pot += ph[(size_t)b * EM_NCAND + l]; }
            if (pot < best) {
                best = pot;
                pick = cand[l];
            }
        }
    }
    HIPCHK(hipMemcpyAsync(dmu.p, mu.data(), sizeof(double) * mu.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dvar.p, var.data(), sizeof(double) * var.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dpi.p, pi.data(), sizeof(double) * pi.size(), hipMemcpyHostToDevice, s));
    HIPCHK(launch_em_lc(dpi.p, dvar.p, K, d, dlc.p, s));
    for (int it = 0; it < em_iters; ++it) { HIPCHK(launch_em_iteration(dx.p, n, d, K, dpi.p, dmu.p, dvar.p, dlc.p, part.p, s)); }
    if (loglik) {
        HIPCHK(launch_em_loglik(dx.p, n, d, K, dmu.p, dvar.p, dlc.p, part.p, s));
        HIPCHK(hipMemcpyAsync(ph.data(), part.p, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipMemcpyAsync(mu_out, dmu.p, sizeof(double) * mu.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(var_out, dvar.p, sizeof(double) * var.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(pi_out, dpi.p, sizeof(double) * pi.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (loglik) {
        double ll = 0.;
        for (int64_t b = 0; b < nb; ++b) { ll += ph[(size_t)b]; }
        *loglik = ll / (double)n;
    }
    return MCI_OK;
}

} // namespace

extern "C" {

// ---- network ----
mci_status mci_mlp_create(int d, int hidden, int ncomp, uint64_t seed, mci_mlp **out)
{
    if (!out) { return fail(MCI_E_INVALID_ARGUMENT, "null output pointer"); }
    *out = nullptr;
    return mlp_make(d, hidden, ncomp, seed, nullptr, out);
}

mci_status mci_mlp_destroy(mci_mlp *m)
{
    delete m;
    return MCI_OK;
}

mci_status mci_mlp_nparams(mci_mlp *m, int64_t *n)
{
    CHECKM(m);
    *n = m->np;
    return MCI_OK;
}

mci_status mci_mlp_get_params(mci_mlp *m, float *p, int64_t n)
{
    CHECKM(m);
    if (n != m->np || !p) { return fail(MCI_E_INVALID_ARGUMENT, "expected %lld parameters", (long long)m->np); }
    std::vector<float> pad((size_t)m->npad);
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(pad.data(), m->p.p, pad.size() * sizeof(float), hipMemcpyDeviceToHost));
    padded_to_flat(m, pad, p);
    return MCI_OK;
}

mci_status mci_mlp_set_params(mci_mlp *m, const float *p, int64_t n)
{
    CHECKM(m);
    if (n != m->np || !p) { return fail(MCI_E_INVALID_ARGUMENT, "expected %lld parameters", (long long)m->np); }
    std::vector<float> pad;
    flat_to_padded(m, p, pad);
    HIPCHK(hipStreamSynchronize(m->stream));
    return mlp_upload_params(m, pad);
}

mci_status mci_mlp_forward(mci_mlp *m, int64_t n, const float *x, float *out)
{
    CHECKM(m);
    if (n < 1 || !x || !out) { return fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    TRY(mlp_reserve(m, n));
    TRY(upload_transposed(m->X.p, x, n, m->d));
    HIPCHK(mcim::launch_to_bf16_scaled(m->X.p, m->Xb.p, (int64_t)m->d * n, mcim::IN_SCALE, m->stream));
    TRY(mlp_forward(m, n, m->Xb.p, m->OUT.p));
    std::vector<float> t((size_t)m->Op * n);
    HIPCHK(hipMemcpyAsync(t.data(), m->OUT.p, t.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (int64_t r = 0; r < n; ++r) {
        for (int c = 0; c < m->O; ++c) { out[(size_t)r * m->O + c] = t[(size_t)c * n + r]; }
    }
    return MCI_OK;
}

mci_status mci_mlp_loss_grad(mci_mlp *m, int64_t n, const float *ctx, const float *y, double *loss, float *grad)
{
    CHECKM(m);
    if (n < 1 || !ctx || !y) { return fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    TRY(mlp_reserve(m, n));
    TRY(upload_transposed(m->X.p, ctx, n, m->d));
    TRY(upload_transposed(m->Y.p, y, n, m->d));
    TRY(mlp_fwd_bwd(m, n));
    std::vector<float> l((size_t)n), gp((size_t)m->npad);
    HIPCHK(hipMemcpyAsync(l.data(), m->loss.p, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(gp.data(), m->g.p, sizeof(float) * m->npad, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0.;
        for (float v : l) { s += v; }
        *loss = s / (double)n;
    }
    if (grad) { padded_to_flat(m, gp, grad); }
    return MCI_OK;
}

// ---- sampler ----
mci_status mci_lp_create(int d, int ncomp, const double *weights, const double *means, const double *covs, int64_t nwalkers, mci_lp **out)
{
    if (!out) { return fail(MCI_E_INVALID_ARGUMENT, "null output pointer"); }
    *out = nullptr;
    if (d < 1 || d > 32 || ncomp < 1 || ncomp > mcil::KMAX || nwalkers < 2 || nwalkers > (int64_t(1) << 30)) {
        return fail(MCI_E_INVALID_ARGUMENT, "need 1 <= d <= 32, 1 <= ncomp <= %d, 2 <= nwalkers", mcil::KMAX);
    }
    if (!weights || !means || !covs) { return fail(MCI_E_INVALID_ARGUMENT, "null parameter array"); }
    mci_lp *lp = new mci_lp();
    lp->d = d;
    lp->K = ncomp;
    lp->W = nwalkers;
    const int KM = mcil::KMAX;
    lp->mu.assign((size_t)KM * d, 0.);
    lp->P.assign((size_t)KM * d * d, 0.);
    lp->U.assign((size_t)KM * d * d, 0.);
    lp->Um.assign((size_t)KM * d, 0.);
    lp->v.assign((size_t)KM * d, 0.);
    lp->L.assign((size_t)KM * d * d, 0.);
    lp->c.assign(KM, -HUGE_VAL);
    lp->cw.assign(KM, 1.);
    double ws = 0.;
    for (int k = 0; k < ncomp; ++k) { ws += weights[k]; }
    if (!(ws > 0.)) {
        delete lp;
        return fail(MCI_E_INVALID_ARGUMENT, "weights must be >= 0 with a positive sum");
    }
    double cum = 0.;
    for (int k = 0; k < ncomp; ++k) {
        if (!(weights[k] >= 0.)) {
            delete lp;
            return fail(MCI_E_INVALID_ARGUMENT, "weights must be >= 0");
        }
        const double *S = covs + (size_t)k * d * d;
        double *Lk = lp->L.data() + (size_t)k * d * d;
        // Cholesky S = L L^T
        for (int i = 0; i < d; ++i) {
            for (int j = 0; j <= i; ++j) {
                double s = S[i * d + j];
                for (int q = 0; q < j; ++q) { s -= Lk[i * d + q] * Lk[j * d + q]; }
                if (i == j) {
                    if (!(s > 0.)) {
                        delete lp;
                        return fail(MCI_E_INVALID_ARGUMENT, "covariance %d is not positive definite", k);
                    }
                    Lk[i * d + i] = std::sqrt(s);
                }
                else { Lk[i * d + j] = s / Lk[j * d + j]; }
            }
        }
        // P = S^-1 from the columns of L^-T L^-1
        double *Pk = lp->P.data() + (size_t)k * d * d;
        std::vector<double> col(d), y(d);
        for (int e = 0; e < d; ++e) {
            for (int i = 0; i < d; ++i) { // L y = e_e
                double s = (i == e) ? 1. : 0.;
                for (int q = 0; q < i; ++q) { s -= Lk[i * d + q] * y[q]; }
                y[i] = s / Lk[i * d + i];
            }
            for (int i = d - 1; i >= 0; --i) { // L^T col = y
                double s = y[i];
                for (int q = i + 1; q < d; ++q) { s -= Lk[q * d + i] * col[q]; }
                col[i] = s / Lk[i * d + i];
            }
            for (int i = 0; i < d; ++i) { Pk[i * d + e] = col[i]; }
        }
        for (int i = 0; i < d; ++i) { // exact symmetry
            for (int j = 0; j < i; ++j) {
                const double a = 0.5 * (Pk[i * d + j] + Pk[j * d + i]);
                Pk[i * d + j] = a;
                Pk[j * d + i] = a;
            }
        }
        // U = (Cholesky factor of P)^T, upper: P = U^T U
        double *Uk = lp->U.data() + (size_t)k * d * d;
        {
            std::vector<double> Lp((size_t)d * d, 0.);
            for (int i = 0; i < d; ++i) {
                for (int j = 0; j <= i; ++j) {
                    double s2 = Pk[i * d + j];
                    for (int q = 0; q < j; ++q) { s2 -= Lp[i * d + q] * Lp[j * d + q]; }
                    if (i == j) { Lp[i * d + i] = std::sqrt(std::fmax(s2, 1e-300)); }
                    else { Lp[i * d + j] = s2 / Lp[j * d + j]; }
                }
            }
            for (int i = 0; i < d; ++i) {
                for (int j = i; j < d; ++j) { Uk[i * d + j] = Lp[j * d + i]; }
            }
            for (int i = 0; i < d; ++i) {
                double sm = 0.;
                for (int j = i; j < d; ++j) { sm += Uk[i * d + j] * means[(size_t)k * d + j]; }
                lp->Um[(size_t)k * d + i] = sm;
            }
        }
        double logdet = 0.;
        for (int i = 0; i < d; ++i) { logdet += 2.0 * std::log(Lk[i * d + i]); }
        const double wn = weights[k] / ws;
        lp->c[k] = std::log(wn) - 0.5 * (logdet + d * std::log(2.0 * M_PI));
        for (int i = 0; i < d; ++i) { lp->mu[(size_t)k * d + i] = means[(size_t)k * d + i]; }
        for (int i = 0; i < d; ++i) {
            double s = 0.;
            for (int j = 0; j < d; ++j) { s += Pk[i * d + j] * means[(size_t)k * d + j]; }
            lp->v[(size_t)k * d + i] = s;
        }
        cum += wn;
        lp->cw[k] = (k == ncomp - 1) ? 1. : cum;
    }
    *out = lp;
    return MCI_OK;
}

mci_status mci_lp_destroy(mci_lp *lp)
{
    delete lp;
    return MCI_OK;
}

mci_status mci_lp_set_seed(mci_lp *lp, uint64_t seed)
{
    CHECKL(lp);
    lp->seed = seed;
    return MCI_OK;
}

mci_status mci_lp_set_network(mci_lp *lp, int hidden, int ncomp)
{
    CHECKL(lp);
    if (hidden < 16 || hidden > 1024 || ncomp < 1 || ncomp > 8) { return fail(MCI_E_INVALID_ARGUMENT, "network sizes out of range"); }
    delete lp->net;
    lp->net = nullptr;
    lp->H = hidden;
    lp->Kq = ncomp;
    return MCI_OK;
}

mci_status mci_lp_run(mci_lp *lp, int mode, int64_t nsteps, int nblocks, double pmix, double *avg, double *err)
{
    CHECKL(lp);
    if (nblocks < 2 || (nblocks & (nblocks - 1)) != 0 || nblocks > 64 || nsteps < nblocks || nsteps % nblocks != 0) {
        return fail(MCI_E_INVALID_ARGUMENT, "nsteps must be a multiple of nblocks, a power of two in 2..64");
    }
    if ((mode == MCI_LP_MIXED_FP32 || mode == MCI_LP_INDEP_FP32) && !(pmix >= 0. && pmix <= 1.)) { return fail(MCI_E_INVALID_ARGUMENT, "mixing probability must be in [0, 1]"); }
    TRY(lp_device(lp));
    switch (mode) {
    case MCI_LP_PLAIN_FP64: TRY(run_plain<double>(lp, nsteps, nblocks)); break;
    case MCI_LP_PLAIN_FP32: TRY(run_plain<float>(lp, nsteps, nblocks)); break;
    case MCI_LP_MIXED_FP32: TRY(run_mixed(lp, nsteps, nblocks, pmix, false)); break;
    case MCI_LP_INDEP_FP32: TRY(run_mixed(lp, nsteps, nblocks, pmix, true)); break;
    default: return fail(MCI_E_DOMAIN, "unknown mode %d", mode);
    }
    if (avg) { avg[0] = lp->results[0], avg[1] = lp->results[1]; }
    if (err) { err[0] = lp->results[2], err[1] = lp->results[3]; }
    return MCI_OK;
}

mci_status mci_lp_collect(mci_lp *lp, int nsnap, int64_t lag)
{
    CHECKL(lp);
    if (nsnap < 2 || lag < 1) { return fail(MCI_E_INVALID_ARGUMENT, "need nsnap >= 2 and lag >= 1"); }
    TRY(lp_device(lp));
    mcil::State<float> S;
    TRY(state_of<float>(lp, 0, 1, S));
    const mcil::Target<float> T = target_of<float>(lp);
    const uint32_t k0 = (uint32_t)lp->seed ^ 0x9E3779B9u, k1 = (uint32_t)(lp->seed >> 32) ^ 0x7F4A7C15u; // streams of their own
    HIPCHK(lp->snap.ensure((size_t)nsnap * lp->d * lp->W));
    HIPCHK(mcil::launch_lp_init<float>(T, S, k0, k1, lp->stream));
    if (lp->step32 <= 0.) { lp->step32 = 1.0; }
    TRY(calibrate<float>(lp, S, lp->step32));
    for (int i = 0; i < nsnap; ++i) {
        HIPCHK(mcil::launch_lp_rw<float>(T, S, (float)lp->step32, lp->t_next, lag, k0, k1, lp->stream));
        lp->t_next += (uint64_t)lag;
        HIPCHK(mcil::launch_lp_snapshot<float>(S, lp->d, lp->snap.p + (size_t)i * lp->d * lp->W, lp->stream));
    }
    HIPCHK(hipStreamSynchronize(lp->stream));
    lp->nsnap = nsnap;
    // mean position of the last snapshot
    std::vector<float> last((size_t)lp->d * lp->W);
    HIPCHK(hipMemcpy(last.data(), lp->snap.p + (size_t)(nsnap - 1) * lp->d * lp->W, last.size() * sizeof(float), hipMemcpyDeviceToHost));
    lp->ctx0.assign(lp->d, 0.f);
    for (int j = 0; j < lp->d; ++j) {
        double sm = 0.;
        for (int64_t w = 0; w < lp->W; ++w) { sm += last[(size_t)j * lp->W + w]; }
        lp->ctx0[j] = (float)(sm / (double)lp->W);
    }
    return MCI_OK;
}

mci_status mci_lp_train(mci_lp *lp, int iters, int batch, double lr, double *loss_first, double *loss_last)
{
    CHECKL(lp);
    if (lp->nsnap < 2) { return fail(MCI_E_RUNTIME, "no training samples: call mci_lp_collect first"); }
    if (iters < 1 || batch < 64 || batch > (1 << 20) || !(lr > 0.)) { return fail(MCI_E_INVALID_ARGUMENT, "bad training settings"); }
    TRY(net_ready(lp));
    mci_mlp *n = lp->net;
    TRY(mlp_reserve(n, batch));
    if (lp->fixed_input) {
        HIPCHK(lp->dctx0.ensure((size_t)lp->d));
        HIPCHK(hipMemcpy(lp->dctx0.p, lp->ctx0.data(), sizeof(float) * (size_t)lp->d, hipMemcpyHostToDevice));
    }
    const uint32_t k0 = (uint32_t)lp->seed ^ 0x85EBCA6Bu, k1 = (uint32_t)(lp->seed >> 32) ^ 0xC2B2AE35u;
    auto t0 = clk::now();
    std::vector<float> l((size_t)batch);
    auto mean_loss = [&](double &out) -> mci_status {
        HIPCHK(hipMemcpyAsync(l.data(), n->loss.p, sizeof(float) * batch, hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
        double s = 0.;
        for (float v : l) { s += v; }
        out = s / batch;
        return MCI_OK;
    };
    for (int it = 0; it < iters; ++it) {
        HIPCHK(mcil::launch_lp_pairs(lp->snap.p, lp->nsnap, lp->d, lp->W, batch, lp->train_iter++, k0, k1, (float)lp->same_frac,
                                     lp->fixed_input ? lp->dctx0.p : nullptr, n->X.p, n->Y.p, n->stream));
        TRY(mlp_fwd_bwd(n, batch));
        if (it == 0) { TRY(mean_loss(lp->loss_first)); }
        TRY(mlp_adam(n, (float)lr));
    }
    TRY(mean_loss(lp->loss_last)); // loss of the last minibatch (before its update)
    lp->timings[3] = ms_since(t0);
    if (loss_first) { *loss_first = lp->loss_first; }
    if (loss_last) { *loss_last = lp->loss_last; }
    return MCI_OK;
}

mci_status mci_lp_set_training_pairs(mci_lp *lp, double same_fraction)
{
    CHECKL(lp);
    if (!(same_fraction >= 0. && same_fraction <= 1.)) { return fail(MCI_E_INVALID_ARGUMENT, "fraction must be in [0, 1]"); }
    lp->same_frac = same_fraction;
    return MCI_OK;
}

// Diagonal Gaussian mixture fitted by EM to x [n][d] (row-major): greedy k-means++ seeds from the
// stream of `seed`, every component starting at the pooled variance with weight 1 / K, then
// em_iters EM iterations (responsibilities by log-sum-exp; variances = E[x^2] - mean^2 + 1e-6).
// Host code only; outputs mu [K][d], var [K][d], pi [K] and the mean log-likelihood.
mci_status mci_gmm_fit_diag(const double *xin, int64_t n, int d, int K, int em_iters, uint64_t seed, double *mu_out, double *var_out,
                            double *pi_out, double *loglik)
{
    if (!xin || !mu_out || !var_out || !pi_out || n < 1 || d < 1 || d > 4096 || K < 1 || K > 64 || n < K) {
        return fail(MCI_E_INVALID_ARGUMENT, "need x, outputs, 1 <= d <= 4096, 1 <= K <= 64 and n >= K");
    }
    if (em_iters < 0 || em_iters > 1000) { return fail(MCI_E_INVALID_ARGUMENT, "EM iterations must be in [0, 1000]"); }
    const int64_t ns = n;
    const double *x = xin;
    // the sums over samples run in NCH fixed chunks on threads of their own and are added in chunk
    // order, so that the fit does not depend on the machine's thread count
    constexpr int NCH = 16;
    auto chunks = [&](auto &&fn) {
        std::vector<std::thread> th;
        for (int c = 0; c < NCH; ++c) { th.emplace_back([&, c] { fn(c, ns * c / NCH, ns * (c + 1) / NCH); }); }
        for (auto &t : th) { t.join(); }
    };
    std::mt19937_64 g(seed ^ 0x5DEECE66DULL);
    std::uniform_real_distribution<double> U(0., 1.);
    std::vector<double> mu((size_t)K * d), var((size_t)K * d), pi((size_t)K, 1.0 / K), d2((size_t)ns, HUGE_VAL);
    double tot = 0.;
    std::vector<double> mean(d, 0.);
    for (int64_t r = 0; r < ns; ++r) {
        for (int j = 0; j < d; ++j) { mean[j] += x[(size_t)r * d + j] / (double)ns; }
    }
    for (int64_t r = 0; r < ns; ++r) {
        for (int j = 0; j < d; ++j) { const double t = x[(size_t)r * d + j] - mean[j]; tot += t * t; }
    }
    const double v0 = std::max(tot / ((double)ns * d), 1e-6);
    // greedy k-means++: each further seed is the best of LTRIALS candidates drawn with probability
    // proportional to the squared distance to the nearest seed (the one that lowers the summed
    // squared distance most), so that well-separated clusters are not left without a seed
    constexpr int LTRIALS = 16;
    auto dist2 = [&](int64_t r, int64_t c) {
        double e = 0.;
        for (int j = 0; j < d; ++j) { const double t = x[(size_t)r * d + j] - x[(size_t)c * d + j]; e += t * t; }
        return e;
    };
    int64_t pick = (int64_t)(U(g) * (double)ns) % ns;
    for (int k = 0; k < K; ++k) {
        for (int j = 0; j < d; ++j) { mu[(size_t)k * d + j] = x[(size_t)pick * d + j]; var[(size_t)k * d + j] = v0; }
        double sum = 0.;
        for (int64_t r = 0; r < ns; ++r) {
            d2[(size_t)r] = std::min(d2[(size_t)r], dist2(r, pick));
            sum += d2[(size_t)r];
        }
        if (k == K - 1) { break; }
        double best = HUGE_VAL;
        for (int l = 0; l < LTRIALS; ++l) {
            const double u = U(g) * sum;
            double c = 0.;
            int64_t cand = ns - 1;
            for (int64_t r = 0; r < ns; ++r) {
                c += d2[(size_t)r];
                if (u < c) { cand = r; break; }
            }
            double part[NCH];
            chunks([&](int c, int64_t r0, int64_t r1) {
                double s = 0.;
                for (int64_t r = r0; r < r1; ++r) { s += std::min(d2[(size_t)r], dist2(r, cand)); }
                part[c] = s;
            });
            double pot = 0.;
            for (int c = 0; c < NCH; ++c) { pot += part[c]; }
            if (pot < best) { best = pot; pick = cand; }
        }
    }
    // EM
    const size_t KS = (size_t)K * (1 + 2 * d); // per chunk: nk [K], s1 [K][d], s2 [K][d]
    std::vector<double> acc(KS * NCH), nk((size_t)K), s1((size_t)K * d), s2((size_t)K * d), lc((size_t)K);
    for (int it = 0; it < em_iters; ++it) {
        for (int k = 0; k < K; ++k) {
            double ld = 0.;
            for (int j = 0; j < d; ++j) { ld += std::log(2. * M_PI * var[(size_t)k * d + j]); }
            lc[(size_t)k] = std::log(pi[(size_t)k]) - 0.5 * ld;
        }
        chunks([&](int c, int64_t r0, int64_t r1) {
            double *a = &acc[KS * c], *ank = a, *as1 = a + K, *as2 = a + K + (size_t)K * d;
            std::fill(a, a + KS, 0.);
            std::vector<double> resp((size_t)K);
            for (int64_t r = r0; r < r1; ++r) {
                const double *xr = &x[(size_t)r * d];
                double mx = -HUGE_VAL;
                for (int k = 0; k < K; ++k) {
                    double q = 0.;
                    for (int j = 0; j < d; ++j) { const double t = xr[j] - mu[(size_t)k * d + j]; q += t * t / var[(size_t)k * d + j]; }
                    resp[(size_t)k] = lc[(size_t)k] - 0.5 * q;
                    mx = std::max(mx, resp[(size_t)k]);
                }
                double z = 0.;
                for (int k = 0; k < K; ++k) { resp[(size_t)k] = std::exp(resp[(size_t)k] - mx); z += resp[(size_t)k]; }
                for (int k = 0; k < K; ++k) {
                    const double rk = resp[(size_t)k] / z;
                    ank[k] += rk;
                    for (int j = 0; j < d; ++j) { as1[(size_t)k * d + j] += rk * xr[j]; as2[(size_t)k * d + j] += rk * xr[j] * xr[j]; }
                }
            }
        });
        std::fill(nk.begin(), nk.end(), 0.);
        std::fill(s1.begin(), s1.end(), 0.);
        std::fill(s2.begin(), s2.end(), 0.);
        for (int c = 0; c < NCH; ++c) {
            const double *a = &acc[KS * c];
            for (int k = 0; k < K; ++k) { nk[(size_t)k] += a[k]; }
            for (size_t e = 0; e < (size_t)K * d; ++e) { s1[e] += a[K + e]; s2[e] += a[K + (size_t)K * d + e]; }
        }
        for (int k = 0; k < K; ++k) {
            const double nkk = nk[(size_t)k] + 1e-9;
            pi[(size_t)k] = std::max(nkk / (double)ns, 1e-12);
            for (int j = 0; j < d; ++j) {
                const double m1 = s1[(size_t)k * d + j] / nkk;
                mu[(size_t)k * d + j] = m1;
                var[(size_t)k * d + j] = std::max(s2[(size_t)k * d + j] / nkk - m1 * m1, 0.) + 1e-6;
            }
        }
    }
    double ll = 0.;
    {
        std::vector<double> lc((size_t)K), q((size_t)K);
        for (int k = 0; k < K; ++k) {
            double ld = 0.;
            for (int j = 0; j < d; ++j) { ld += std::log(2. * M_PI * var[(size_t)k * d + j]); }
            lc[(size_t)k] = std::log(pi[(size_t)k]) - 0.5 * ld;
        }
        for (int64_t r = 0; r < ns; ++r) {
            double mx = -HUGE_VAL;
            for (int k = 0; k < K; ++k) {
                double e = 0.;
                for (int j = 0; j < d; ++j) { const double t = x[(size_t)r * d + j] - mu[(size_t)k * d + j]; e += t * t / var[(size_t)k * d + j]; }
                q[(size_t)k] = lc[(size_t)k] - 0.5 * e;
                mx = std::max(mx, q[(size_t)k]);
            }
            double z = 0.;
            for (int k = 0; k < K; ++k) { z += std::exp(q[(size_t)k] - mx); }
            ll += mx + std::log(z);
        }
    }
    std::copy(mu.begin(), mu.end(), mu_out);
    std::copy(var.begin(), var.end(), var_out);
    std::copy(pi.begin(), pi.end(), pi_out);
    if (loglik) { *loglik = ll / (double)ns; }
    return MCI_OK;
}

// The network's last layer started at the diagonal mixture that EM fits (mci_gmm_fit_diag, seeded
// from the sampler's seed) to the first 16,384 walkers of the last snapshot: biases = (means / 4,
// log-scales = log sqrt(variances) clamped inside [-3, 3], log-weights), weights scaled by 0.02,
// so that before training the network's mixture is about that fit at every input.  Training
// then refines the whole network.
mci_status mci_gmm_fit_diag_device(const double *x, int64_t n, int d, int K, int em_iters, uint64_t seed, double *mu, double *var,
                                   double *pi, double *loglik)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { return fail(MCI_E_DEVICE, "no GPU visible"); }
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const mci_status st = gmm_fit_device(x, n, d, K, em_iters, seed, s, mu, var, pi, loglik);
    (void)hipStreamDestroy(s);
    return st;
}

mci_status mci_lp_init_network_from_samples(mci_lp *lp, int em_iters)
{
    CHECKL(lp);
    if (lp->nsnap < 1) { return fail(MCI_E_RUNTIME, "no training samples: call mci_lp_collect first"); }
    if (em_iters < 1 || em_iters > 1000) { return fail(MCI_E_INVALID_ARGUMENT, "EM iterations must be in [1, 1000]"); }
    TRY(net_ready(lp));
    mci_mlp *n = lp->net;
    const int d = lp->d, K = n->K, H = n->H, O = n->O;
    const int64_t W = lp->W, ns = std::min<int64_t>(W, 16384);
    if (ns < K) { return fail(MCI_E_INVALID_ARGUMENT, "fewer walkers than mixture components"); }
    std::vector<float> last((size_t)d * W);
    HIPCHK(hipStreamSynchronize(lp->stream));
    HIPCHK(hipMemcpy(last.data(), lp->snap.p + (size_t)(lp->nsnap - 1) * d * W, last.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<double> x((size_t)ns * d);
    for (int64_t r = 0; r < ns; ++r) {
        for (int j = 0; j < d; ++j) { x[(size_t)r * d + j] = last[(size_t)j * W + r]; }
    }
    std::vector<double> mu((size_t)K * d), var((size_t)K * d), pi((size_t)K);
    TRY(gmm_fit_device(x.data(), ns, d, K, em_iters, lp->seed, lp->stream, mu.data(), var.data(), pi.data(), nullptr));
    // into the last layer
    std::vector<float> p((size_t)n->np);
    TRY(mci_mlp_get_params(n, p.data(), n->np));
    const int64_t oW3 = (int64_t)d * H + H + (int64_t)H * H + H, ob3 = oW3 + (int64_t)H * O;
    for (int64_t e = oW3; e < ob3; ++e) { p[(size_t)e] *= 0.02f; }
    for (int k = 0; k < K; ++k) {
        for (int j = 0; j < d; ++j) {
            p[(size_t)(ob3 + 2 * d * k + j)] = (float)(mu[(size_t)k * d + j] / mcim::MU_SCALE);
            const double ls = 0.5 * std::log(var[(size_t)k * d + j]);
            p[(size_t)(ob3 + 2 * d * k + d + j)] = (float)std::fmin(std::fmax(ls, -(double)mcim::LS_MAX + 1e-3), (double)mcim::LS_MAX - 1e-3);
        }
        p[(size_t)(ob3 + 2 * d * K + k)] = (float)std::log(pi[(size_t)k]);
    }
    TRY(mci_mlp_set_params(n, p.data(), n->np));
    return MCI_OK;
}

mci_status mci_lp_set_fixed_input(mci_lp *lp, int on)
{
    CHECKL(lp);
    lp->fixed_input = on != 0;
    return MCI_OK;
}

mci_status mci_lp_get_results(mci_lp *lp, double *r, int n)
{
    CHECKL(lp);
    for (int i = 0; i < n && i < 8; ++i) { r[i] = lp->results[i]; }
    return MCI_OK;
}

mci_status mci_lp_get_timings(mci_lp *lp, double *ms, int n)
{
    CHECKL(lp);
    for (int i = 0; i < n && i < 4; ++i) { ms[i] = lp->timings[i]; }
    return MCI_OK;
}

mci_status mci_lp_get_walker_means(mci_lp *lp, double *m, int64_t n)
{
    CHECKL(lp);
    if (n != 2 * lp->W || lp->h_wavg.size() != (size_t)n) { return fail(MCI_E_INVALID_ARGUMENT, "expected 2 x nwalkers values after a run"); }
    std::copy(lp->h_wavg.begin(), lp->h_wavg.end(), m);
    return MCI_OK;
}

mci_status mci_lp_get_network(mci_lp *lp, mci_mlp **m)
{
    CHECKL(lp);
    TRY(lp_device(lp));
    TRY(net_ready(lp));
    *m = lp->net;
    return MCI_OK;
}

mci_status mci_lp_get_mixture(mci_lp *lp, int kcap, int *K, double *weights, double *means, double *sigmas)
{
    CHECKL(lp);
    if (!K || !weights || !means || !sigmas) { return fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    TRY(lp_device(lp));
    TRY(net_ready(lp));
    mcil::NetDev N;
    TRY(net_dev(lp, N));
    if (kcap < N.K) { return fail(MCI_E_INVALID_ARGUMENT, "room for %d components needed", N.K); }
    mcil::IndepMix M{};
    std::vector<double> w;
    TRY(indep_mixture(lp, N, M, &w));
    const int d = lp->d;
    const size_t KD = (size_t)mcil::KMAX * d;
    std::vector<float> h(3 * KD);
    HIPCHK(hipMemcpy(h.data(), lp->imix.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < N.K; ++k) {
        weights[k] = w[(size_t)k];
        for (int j = 0; j < d; ++j) {
            means[(size_t)k * d + j] = (double)h[(size_t)k * d + j];
            sigmas[(size_t)k * d + j] = (double)h[2 * KD + (size_t)k * d + j];
        }
    }
    *K = N.K;
    return MCI_OK;
}

mci_status mci_lp_indep_move(mci_lp *lp, int64_t n, float *x, int nl, uint64_t t0, int32_t *accepted, float *mix)
{
    CHECKL(lp);
    if (n < 1 || !x || !accepted || nl < 1) { return fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    TRY(lp_device(lp));
    TRY(net_ready(lp));
    mcil::NetDev N;
    TRY(net_dev(lp, N));
    mcil::IndepMix M{};
    TRY(indep_mixture(lp, N, M));
    const int d = lp->d;
    const size_t W = (size_t)n;
    DBuf<float> xs, q, m, S, sb, cb;
    DBuf<double> sq, bm;
    DBuf<int> jb, bi;
    DBuf<int32_t> na, nai;
    HIPCHK(xs.ensure(W * d));
    HIPCHK(q.ensure(W * mcil::KMAX));
    HIPCHK(m.ensure(W));
    HIPCHK(S.ensure(W));
    HIPCHK(sb.ensure(2 * W));
    HIPCHK(cb.ensure(2 * W));
    HIPCHK(sq.ensure(W));
    HIPCHK(bm.ensure(2 * W));
    HIPCHK(jb.ensure(W));
    HIPCHK(bi.ensure(W));
    HIPCHK(na.ensure(W));
    HIPCHK(nai.ensure(W));
    HIPCHK(hipMemcpy(xs.p, x, sizeof(float) * W * d, hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(nai.p, 0, sizeof(int32_t) * W, lp->stream));
    mcil::State<float> St{n, xs.p, q.p, m.p, S.p, sq.p, sb.p, cb.p, jb.p, bi.p, bm.p, 0, na.p};
    const mcil::Target<float> T = target_of<float>(lp);
    HIPCHK(mcil::launch_lp_state(T, St, lp->stream));
    HIPCHK(mcil::launch_lp_indep(M, T, St, t0, nl, (uint32_t)lp->seed, (uint32_t)(lp->seed >> 32), nai.p, lp->stream));
    HIPCHK(hipStreamSynchronize(lp->stream));
    HIPCHK(hipMemcpy(x, xs.p, sizeof(float) * W * d, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(accepted, nai.p, sizeof(int32_t) * W, hipMemcpyDeviceToHost));
    if (mix) {
        const size_t KD = (size_t)mcil::KMAX * d;
        HIPCHK(hipMemcpy(mix, lp->imix.p, sizeof(float) * (3 * KD + 2 * mcil::KMAX), hipMemcpyDeviceToHost));
    }
    return MCI_OK;
}

mci_status mci_lp_eval_network(mci_lp *lp, int64_t n, const float *x, float *out)
{
    CHECKL(lp);
    if (n < 1 || n > (int64_t)1 << 24 || !x || !out) { return fail(MCI_E_INVALID_ARGUMENT, "need 1 <= n <= 2^24 and x, out"); }
    TRY(lp_device(lp));
    mcil::NetDev N;
    TRY(net_dev(lp, N));
    DBuf<float> dx, dout;
    HIPCHK(dx.ensure((size_t)lp->d * n));
    HIPCHK(dout.ensure((size_t)N.O * n));
    HIPCHK(hipMemcpy(dx.p, x, sizeof(float) * (size_t)lp->d * n, hipMemcpyHostToDevice));
    mcil::State<float> S{};
    S.W = n;
    HIPCHK(mcil::launch_lp_network(N, target_of<float>(lp), S, dx.p, dout.p, lp->stream));
    HIPCHK(hipStreamSynchronize(lp->stream));
    HIPCHK(hipMemcpy(out, dout.p, sizeof(float) * (size_t)N.O * n, hipMemcpyDeviceToHost));
    return MCI_OK;
}

} // extern "C"
