"""Device functor sources of the tests, written as a caller writes them for
mci_add_sampling_function_device / mci_add_observable_device (include/mci_amd/device_functor.hpp):

- the reference harness's plugins outside the library's catalogue (oracle/ref_harness.cpp
  QuarticPDF, MomentsObs), whose chains the reference wrote into tests/golden/device_functors.json;
- restatements of the catalogue's diagonal Gaussian and x_i^2, which must give the catalogue's
  chains bit for bit (all-coordinate moves recompute every proto value in both)."""

PDF_USER_QUARTIC = 101   # the harness's kind codes
PDF_USER_QUARTIC_UPD = 102
OBS_USER_MOMENTS = 101

QUARTIC = r"""
__device__ void quartic_proto(int ndim, MciVec x, const double *par, MciVec proto)
{
    for (int i = 0; i < ndim; ++i) {
        const double q = x[i] * x[i];
        proto[i] = (q * q) * par[i];
    }
}
__device__ double quartic_acceptance(int nproto, MciVec po, MciVec pn, const double *par)
{
    double e = 0.0;
    for (int i = 0; i < nproto; ++i) { e = e + po[i]; }
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pn[i]; }
    return exp(e - s);
}
__device__ double quartic_sampling(int nproto, MciVec pv, const double *par)
{
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pv[i]; }
    return exp(-s);
}
"""

# QuarticUpdPDF: the same with an updatedAcceptance of its own (the changed coordinates only)
QUARTIC_UPD = QUARTIC.replace("quartic_", "quarticupd_") + r"""
__device__ double quarticupd_updated_acceptance(int ndim, int nproto, MciVec xnew, MciChanged changed, MciVec po, MciVec pn,
                                                const double *par)
{
    double e = 0.0;
    for (int i = 0; i < changed.n; ++i) {
        const int j = changed[i];
        const double q = xnew[j] * xnew[j];
        pn[j] = (q * q) * par[j];
        e += pn[j] - po[j];
    }
    return exp(-e);
}
"""

MOMENTS = r"""
__device__ void moments_observable(int ndim, MciVec x, const double *par, MciVec out)
{
    double s = 0.0, m = 0.0;
    for (int i = 0; i < ndim; ++i) {
        const double a = fabs(x[i]);
        s = s + (a * a) * a;
        m = (a > m) ? a : m;
    }
    out[0] = s;
    out[1] = m;
}
"""

DIAG = r"""
__device__ void diag_proto(int ndim, MciVec x, const double *par, MciVec proto)
{
    for (int i = 0; i < ndim; ++i) { const double v = x[i]; proto[i] = (v * v) * par[i]; }
}
__device__ double diag_acceptance(int nproto, MciVec po, MciVec pn, const double *par)
{
    double e = 0.0;
    for (int i = 0; i < nproto; ++i) { e = e + po[i]; }
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pn[i]; }
    return exp(e - s);
}
__device__ double diag_sampling(int nproto, MciVec pv, const double *par)
{
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pv[i]; }
    return exp(-s);
}
"""

X2 = r"""
__device__ void xsq_observable(int ndim, MciVec x, const double *par, MciVec out)
{
    for (int i = 0; i < ndim; ++i) { const double v = x[i]; out[i] = v * v; }
}
"""

# trial moves outside the SRRD family (oracle/ref_harness.cpp PairMove, ScaleMove)
MOVE_USER_PAIR = 201
MOVE_USER_SCALE = 202

PAIR = r"""
__device__ double pair_move(int ndim, MciVec x, MciMoved &moved, MciVec po, MciVec pn, MciVec steps, MciRng &rng, const double *par)
{
    int i = (int)(mci_uniform(rng) * ndim);
    if (i > ndim - 1) { i = ndim - 1; }
    const int n = (i + 1 < ndim) ? 2 : 1;
    for (int k = 0; k < n; ++k) {
        const int j = i + k;
        const double r = 2.0 * mci_uniform(rng) - 1.0;
        x[j] = x[j] + steps[j < ndim / 2 ? 0 : 1] * r;
    }
    moved.first = i;
    moved.n = n;
    return 1.0;
}
"""

SCALE = r"""
__device__ void scale_proto(int ndim, MciVec x, const double *par, MciVec proto)
{
    double q = 0.0;
    for (int i = 0; i < ndim; ++i) { q = q + x[i] * x[i]; }
    proto[0] = q;
}
__device__ double scale_move(int ndim, MciVec x, MciMoved &moved, MciVec po, MciVec pn, MciVec steps, MciRng &rng, const double *par)
{
    const double s = 1.0 + steps[0] * (2.0 * mci_uniform(rng) - 1.0);
    for (int i = 0; i < ndim; ++i) { x[i] = x[i] * s; }
    moved.first = 0;
    moved.n = ndim;
    pn[0] = (s * s) * po[0];
    return pow(pn[0] / po[0], 0.5 * ndim);
}
"""

# kind code -> (name, source, number of proto values or observable values for ndim)
PDFS = {PDF_USER_QUARTIC: ("quartic", QUARTIC, lambda d: d), PDF_USER_QUARTIC_UPD: ("quarticupd", QUARTIC_UPD, lambda d: d)}
OBS = {OBS_USER_MOMENTS: ("moments", MOMENTS, lambda d: 2)}
# kind code -> (name, source, nproto, typeEnds for ndim, change rate for ndim, initial steps from the params)
MOVES = {MOVE_USER_PAIR: ("pair", PAIR, 0, lambda d: [d // 2, d], lambda d: 2.0 / d, lambda p: list(p)),
         MOVE_USER_SCALE: ("scale", SCALE, 1, lambda d: None, lambda d: 1.0, lambda p: [p[0]])}
