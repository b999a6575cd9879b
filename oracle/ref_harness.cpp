// ref_harness.cpp — drives the REFERENCE library (built from its own sources by
// oracle/Makefile into oracle/_ref/) to write golden vectors and to time the
// reference CPU path.  TEST INFRASTRUCTURE ONLY; never part of the product.
//
// It uses only the reference's public plugin API (mci::MCI, SamplingFunctionInterface,
// ObservableFunctionInterface, estimator functions, accumulator classes).  The
// sampling functions and observables below are this project's functor catalogue
// (include/mci_amd.h) written as reference plugins with the same expressions, so
// the reference computes exactly what the device functors compute.
//
// Input: a line-oriented script on stdin.  Output: one JSON object per result line;
// doubles are printed as C99 hex-floats ("%a") so they round-trip bit-exactly.
#include "mci/MCIntegrator.hpp"
#include "mci/DependentObservableInterface.hpp"
#include "mci/Factories.hpp"
#include "mci/MultiStepMove.hpp"
#include "mci/Estimators.hpp"
#include "mci/BlockAccumulator.hpp"
#include "mci/FullAccumulator.hpp"
#include "mci/SimpleAccumulator.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace
{
// functor kinds, as in include/mci_amd.h
enum { PDF_GAUSS_ISO = 1, PDF_GAUSS_DIAG = 2, PDF_EXP_DIAG = 3, PDF_GMM_ISO = 4, PDF_GMM_FULL = 5 };
enum { OBS_X2 = 1, OBS_X2_SUM = 2, OBS_X = 3, OBS_X_SUM = 4, OBS_CONST = 5, OBS_EXP_X2 = 6, OBS_INDICATOR = 7, OBS_GAUSS_X2 = 8,
       OBS_PDF_LOG = 9, OBS_PDF_PROTO = 10, OBS_DEP_SQUARE = 11 };
// plugins outside the library's catalogue, written as a reference user writes them (the fixtures of
// the library's device functors, which take them as source)
enum { PDF_USER_QUARTIC = 101, PDF_USER_QUARTIC_UPD = 102, OBS_USER_MOMENTS = 101 };
enum { MOVE_USER_PAIR = 201, MOVE_USER_SCALE = 202 };

std::string hx(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "\"%a\"", v);
    return b;
}

std::string hxv(const double *v, int n)
{
    std::string s = "[";
    for (int i = 0; i < n; ++i) { if (i) s += ","; s += hx(v[i]); }
    return s + "]";
}

// ---- sampling functions ----
// Every sampling function of the harness keeps what its observation callback last received
// (SamplingFunctionInterface::observationCallback, SamplingFunctionInterface.hpp:100-104) for the
// dependent observables below, and the log of its value in the form of its samplingFunction.
class ObservedPDF: public mci::SamplingFunctionInterface
{
public:
    std::vector<double> seen_pv;  // proto values at the last callback
    int64_t ncallbacks = 0;
    ObservedPDF(int nd, int np): mci::SamplingFunctionInterface(nd, np) {}
    void observationCallback(const double /*x*/[], const double pv[]) override
    {
        seen_pv.assign(pv, pv + _nproto);
        ++ncallbacks;
    }
    virtual double logValue(const double pv[]) const = 0;
};

class GaussIso final: public ObservedPDF
{
protected:
    mci::SamplingFunctionInterface *_clone() const final { return new GaussIso(_ndim); }
public:
    explicit GaussIso(int nd): ObservedPDF(nd, 1) {}
    double logValue(const double pv[]) const final { double s = 0.; s = s + pv[0]; return -s; }
    void protoFunction(const double in[], double pv[]) final
    {
        double s = in[0]*in[0];
        for (int i = 1; i < _ndim; ++i) { s = s + in[i]*in[i]; }
        pv[0] = s;
    }
    double samplingFunction(const double pv[]) const final { return exp(-pv[0]); }
    double acceptanceFunction(const double po[], const double pn[]) const final { return exp(-pn[0] + po[0]); }
};

class DiagPDF final: public ObservedPDF
{
    bool _abs;
    std::vector<double> _a;
protected:
    mci::SamplingFunctionInterface *_clone() const final { return new DiagPDF(_ndim, _abs, _a); }
public:
    DiagPDF(int nd, bool absform, std::vector<double> a): ObservedPDF(nd, nd), _abs(absform), _a(std::move(a))
    {
        if (_a.empty()) { _a.assign(nd, 1.); }
    }
    double logValue(const double pv[]) const final { return -std::accumulate(pv, pv + _nproto, 0.); }
    double f(double x, int i) const { return _abs ? fabs(x)*_a[i] : (x*x)*_a[i]; }
    void protoFunction(const double in[], double pv[]) final
    {
        for (int i = 0; i < _ndim; ++i) { pv[i] = f(in[i], i); }
    }
    double samplingFunction(const double pv[]) const final { return exp(-std::accumulate(pv, pv + _nproto, 0.)); }
    double acceptanceFunction(const double po[], const double pn[]) const final
    {
        double e = std::accumulate(po, po + _nproto, 0.);
        e -= std::accumulate(pn, pn + _nproto, 0.);
        return exp(e);
    }
    double updatedAcceptance(const mci::WalkerState &wlk, const double po[], double pn[]) final
    {
        double e = 0.;
        for (int i = 0; i < wlk.nchanged; ++i) {
            const int j = wlk.changedIdx[i];
            pn[j] = f(wlk.xnew[j], j);
            e += pn[j] - po[j];
        }
        return exp(-e);
    }
};

// p(x) ~ exp(-sum_i c_i x_i^4), one proto value per coordinate; without an updatedAcceptance of its
// own (the interface's default: full recomputation, SamplingFunctionInterface.hpp:85-90) or, as
// QuarticUpdPDF, with one that updates the changed coordinates' proto values
class QuarticPDF: public ObservedPDF
{
protected:
    std::vector<double> _c;
    mci::SamplingFunctionInterface *_clone() const override { return new QuarticPDF(_ndim, _c); }
public:
    QuarticPDF(int nd, std::vector<double> c): ObservedPDF(nd, nd), _c(std::move(c)) {}
    double logValue(const double pv[]) const final { return -std::accumulate(pv, pv + _nproto, 0.); }
    void protoFunction(const double in[], double pv[]) final
    {
        for (int i = 0; i < _ndim; ++i) {
            const double q = in[i]*in[i];
            pv[i] = (q*q)*_c[i];
        }
    }
    double samplingFunction(const double pv[]) const final { return exp(-std::accumulate(pv, pv + _nproto, 0.)); }
    double acceptanceFunction(const double po[], const double pn[]) const final
    {
        double e = std::accumulate(po, po + _nproto, 0.);
        e -= std::accumulate(pn, pn + _nproto, 0.);
        return exp(e);
    }
};

class QuarticUpdPDF final: public QuarticPDF
{
protected:
    mci::SamplingFunctionInterface *_clone() const final { return new QuarticUpdPDF(_ndim, _c); }
public:
    QuarticUpdPDF(int nd, std::vector<double> c): QuarticPDF(nd, std::move(c)) {}
    double updatedAcceptance(const mci::WalkerState &wlk, const double po[], double pn[]) final
    {
        double e = 0.;
        for (int i = 0; i < wlk.nchanged; ++i) {
            const int j = wlk.changedIdx[i];
            const double q = wlk.xnew[j]*wlk.xnew[j];
            pn[j] = (q*q)*_c[j];
            e += pn[j] - po[j];
        }
        return exp(-e);
    }
};

class GmmIso final: public ObservedPDF
{
    int _K;
    std::vector<double> _w, _s, _mu, _c, _b;
protected:
    mci::SamplingFunctionInterface *_clone() const final { return new GmmIso(_ndim, _K, _w, _s, _mu); }
public:
    double logValue(const double pv[]) const final { return pv[_K]; }
    GmmIso(int nd, int K, std::vector<double> w, std::vector<double> s, std::vector<double> mu):
            ObservedPDF(nd, K + 1), _K(K), _w(std::move(w)), _s(std::move(s)), _mu(std::move(mu))
    {
        for (int k = 0; k < _K; ++k) {
            _c.push_back(log(_w[k]) - static_cast<double>(_ndim)*log(_s[k]));
            _b.push_back(0.5/(_s[k]*_s[k]));
        }
    }
    double logp(const double r2[]) const
    {
        double m = -std::numeric_limits<double>::infinity();
        for (int k = 0; k < _K; ++k) { const double t = _c[k] - _b[k]*r2[k]; if (t > m) { m = t; } }
        double s = 0.;
        for (int k = 0; k < _K; ++k) { s = s + exp((_c[k] - _b[k]*r2[k]) - m); }
        return m + log(s);
    }
    void protoFunction(const double in[], double pv[]) final
    {
        for (int k = 0; k < _K; ++k) {
            double s = 0.;
            for (int i = 0; i < _ndim; ++i) { const double d = in[i] - _mu[k*_ndim + i]; s = s + d*d; }
            pv[k] = s;
        }
        pv[_K] = logp(pv);
    }
    double samplingFunction(const double pv[]) const final { return exp(pv[_K]); }
    double acceptanceFunction(const double po[], const double pn[]) const final { return exp(pn[_K] - po[_K]); }
    double updatedAcceptance(const mci::WalkerState &wlk, const double po[], double pn[]) final
    {
        for (int k = 0; k < _K; ++k) {
            double dr = 0.;
            for (int i = 0; i < wlk.nchanged; ++i) {
                const int j = wlk.changedIdx[i];
                const double dn = wlk.xnew[j] - _mu[k*_ndim + j], dd = wlk.xold[j] - _mu[k*_ndim + j];
                dr = dr + (dn*dn - dd*dd);
            }
            pn[k] = po[k] + dr;
        }
        pn[_K] = logp(pn);
        return exp(pn[_K] - po[_K]);
    }
};

// mixture of K full-covariance Gaussians held as precision matrices P_k and
// c_k = log w_k - log det(2 pi Sigma_k) / 2; protovalues q_k = (x - mu_k)^T P_k (x - mu_k) and log p.
// A move of the coordinates in C changes q_k by 2 sum_{a in C} delta_a (P_k (x - mu_k))_a
// + sum_{a, b in C} delta_a delta_b (P_k)_ab.  L_k (Cholesky factor of Sigma_k) draws exact starts.
class GmmFull final: public ObservedPDF
{
    int _K;
    std::vector<double> _w, _mu, _P, _c, _L;
protected:
    mci::SamplingFunctionInterface *_clone() const final { return new GmmFull(_ndim, _K, _w, _mu, _P, _c, _L); }
public:
    GmmFull(int nd, int K, std::vector<double> w, std::vector<double> mu, std::vector<double> P, std::vector<double> c, std::vector<double> L):
            ObservedPDF(nd, K + 1), _K(K), _w(std::move(w)), _mu(std::move(mu)), _P(std::move(P)), _c(std::move(c)),
            _L(std::move(L)) {}
    double logValue(const double pv[]) const final { return pv[_K]; }
    double logp(const double q[]) const
    {
        double m = -std::numeric_limits<double>::infinity();
        for (int k = 0; k < _K; ++k) { const double t = _c[k] - 0.5*q[k]; if (t > m) { m = t; } }
        double s = 0.;
        for (int k = 0; k < _K; ++k) { s = s + exp((_c[k] - 0.5*q[k]) - m); }
        return m + log(s);
    }
    void protoFunction(const double in[], double pv[]) final
    {
        for (int k = 0; k < _K; ++k) {
            const double *P = &_P[(size_t)k*_ndim*_ndim], *mu = &_mu[(size_t)k*_ndim];
            double q = 0.;
            for (int i = 0; i < _ndim; ++i) {
                double r = 0.;
                for (int j = 0; j < _ndim; ++j) { r = r + P[i*_ndim + j]*(in[j] - mu[j]); }
                q = q + (in[i] - mu[i])*r;
            }
            pv[k] = q;
        }
        pv[_K] = logp(pv);
    }
    double samplingFunction(const double pv[]) const final { return exp(pv[_K]); }
    double acceptanceFunction(const double po[], const double pn[]) const final { return exp(pn[_K] - po[_K]); }
    double updatedAcceptance(const mci::WalkerState &wlk, const double po[], double pn[]) final
    {
        for (int k = 0; k < _K; ++k) {
            const double *P = &_P[(size_t)k*_ndim*_ndim], *mu = &_mu[(size_t)k*_ndim];
            double dq = 0.;
            for (int a = 0; a < wlk.nchanged; ++a) {
                const int ja = wlk.changedIdx[a];
                const double da = wlk.xnew[ja] - wlk.xold[ja];
                double r = 0.;
                for (int i = 0; i < _ndim; ++i) { r = r + P[ja*_ndim + i]*(wlk.xold[i] - mu[i]); }
                dq = dq + 2.*da*r;
                for (int b = 0; b < wlk.nchanged; ++b) {
                    const int jb = wlk.changedIdx[b];
                    dq = dq + da*(wlk.xnew[jb] - wlk.xold[jb])*P[ja*_ndim + jb];
                }
            }
            pn[k] = po[k] + dq;
        }
        pn[_K] = logp(pn);
        return exp(pn[_K] - po[_K]);
    }
    // an exact draw of p: component by weight, x = mu_k + L_k z
    void draw(std::mt19937_64 &g, double x[]) const
    {
        std::uniform_real_distribution<double> u(0., 1.);
        std::normal_distribution<double> n(0., 1.);
        double c = u(g), cum = 0.;
        int k = _K - 1;
        for (int i = 0; i < _K - 1; ++i) { cum += _w[i]; if (c < cum) { k = i; break; } }
        std::vector<double> z(_ndim);
        for (int i = 0; i < _ndim; ++i) { z[i] = n(g); }
        for (int i = 0; i < _ndim; ++i) {
            double v = _mu[(size_t)k*_ndim + i];
            for (int j = 0; j <= i; ++j) { v += _L[(size_t)k*_ndim*_ndim + i*_ndim + j]*z[j]; }
            x[i] = v;
        }
    }
};

// ---- observables ----
class CatObs final: public mci::ObservableFunctionInterface
{
    int _kind;
    std::vector<double> _p;
    static int nobsOf(int kind, int nd) { return (kind == OBS_X2 || kind == OBS_X) ? nd : 1; }
protected:
    mci::ObservableFunctionInterface *_clone() const final { return new CatObs(_ndim, _kind, _p); }
public:
    CatObs(int nd, int kind, std::vector<double> p):
            mci::ObservableFunctionInterface(nd, nobsOf(kind, nd), kind == OBS_X2), _kind(kind), _p(std::move(p)) {}
    double par(size_t i, double d) const { return _p.size() > i ? _p[i] : d; }
    void observableFunction(const double in[], double out[]) final
    {
        switch (_kind) {
        case OBS_X2: for (int i = 0; i < _ndim; ++i) { out[i] = in[i]*in[i]; } break;
        case OBS_X2_SUM: { double s = 0.; for (int i = 0; i < _ndim; ++i) { s = s + in[i]*in[i]; } out[0] = s/par(0, 1.); break; }
        case OBS_X: for (int i = 0; i < _ndim; ++i) { out[i] = in[i]; } break;
        case OBS_X_SUM: { double s = 0.; for (int i = 0; i < _ndim; ++i) { s = s + in[i]; } out[0] = s; break; }
        case OBS_CONST: out[0] = par(0, 1.); break;
        case OBS_EXP_X2: { double s = 0.; for (int i = 0; i < _ndim; ++i) { s = s + in[i]*in[i]; } out[0] = exp(-(s/par(0, _ndim))); break; }
        case OBS_INDICATOR: out[0] = (in[static_cast<int>(par(0, 0.))] > 0.) ? 1. : 0.; break;
        case OBS_GAUSS_X2: { double s = 0.; for (int i = 0; i < _ndim; ++i) { s = s + in[i]*in[i]; } out[0] = exp(-s)*s*par(0, 1./sqrt(M_PI*M_PI*M_PI)/3.); break; }
        default: throw std::invalid_argument("unknown observable kind");
        }
    }
    void updatedObservable(const double in[], const int /*nchanged*/, const bool flags[], double out[]) final
    {
        if (_kind == OBS_X2) {
            for (int i = 0; i < _ndim; ++i) { if (flags[i]) { out[i] = in[i]*in[i]; } }
        }
        else { observableFunction(in, out); }
    }
};

// trial moves outside the SRRD family, written as a reference user writes them (TrialMoveInterface,
// drawing from MCI's generator through a uniform_real_distribution(0, 1)).
// PairMove: a uniformly chosen coordinate i and, if it exists, i + 1, each moved by its step type's
// size times a uniform on (-1, 1); step type 0 covers [0, ndim / 2), type 1 the rest; factor 1.
class PairMove final: public mci::TrialMoveInterface
{
    std::vector<double> _s;
    std::uniform_real_distribution<double> _rd{0., 1.};
protected:
    mci::TrialMoveInterface *_clone() const final { return new PairMove(_ndim, _s); }
public:
    PairMove(int nd, std::vector<double> s): mci::TrialMoveInterface(nd, 0), _s(std::move(s)) {}
    int getNStepSizes() const final { return 2; }
    double getStepSize(int i) const final { return _s[i]; }
    void setStepSize(int i, double v) final { _s[i] = v; }
    double getChangeRate() const final { return 2. / _ndim; }
    int getStepSizeIndex(int xidx) const final { return xidx < _ndim / 2 ? 0 : 1; }
    void protoFunction(const double /*in*/[], double /*out*/[]) final {}
    double trialMove(mci::WalkerState &wlk, const double /*po*/[], double /*pn*/[]) final
    {
        int i = static_cast<int>(_rd(*_rgen) * _ndim);
        if (i > _ndim - 1) { i = _ndim - 1; }
        const int n = (i + 1 < _ndim) ? 2 : 1;
        for (int k = 0; k < n; ++k) {
            const int j = i + k;
            const double r = 2. * _rd(*_rgen) - 1.;
            wlk.xnew[j] = wlk.xnew[j] + _s[getStepSizeIndex(j)] * r;
            wlk.changedIdx[k] = j;
        }
        wlk.nchanged = n;
        return 1.;
    }
};

// ScaleMove: every coordinate times s = 1 + step (2u - 1), with one proto value |x|^2 (updated as
// s^2 |x|^2) and the move factor (protonew / protoold)^(ndim / 2) from the proto values: a
// non-symmetric move whose factor depends on the move's own proto bookkeeping
class ScaleMove final: public mci::TrialMoveInterface
{
    double _s;
    std::uniform_real_distribution<double> _rd{0., 1.};
protected:
    mci::TrialMoveInterface *_clone() const final { return new ScaleMove(_ndim, _s); }
public:
    ScaleMove(int nd, double s): mci::TrialMoveInterface(nd, 1), _s(s) {}
    int getNStepSizes() const final { return 1; }
    double getStepSize(int /*i*/) const final { return _s; }
    void setStepSize(int /*i*/, double v) final { _s = v; }
    double getChangeRate() const final { return 1.; }
    int getStepSizeIndex(int /*xidx*/) const final { return 0; }
    void protoFunction(const double in[], double out[]) final
    {
        double q = 0.;
        for (int i = 0; i < _ndim; ++i) { q = q + in[i]*in[i]; }
        out[0] = q;
    }
    double trialMove(mci::WalkerState &wlk, const double po[], double pn[]) final
    {
        const double s = 1. + _s * (2. * _rd(*_rgen) - 1.);
        for (int i = 0; i < _ndim; ++i) {
            wlk.xnew[i] = wlk.xnew[i] * s;
            wlk.changedIdx[i] = i;
        }
        wlk.nchanged = _ndim;
        pn[0] = (s*s)*po[0];
        return pow(pn[0]/po[0], 0.5*_ndim);
    }
};

// sum_i |x_i|^3 and max_i |x_i|, not updateable
class MomentsObs final: public mci::ObservableFunctionInterface
{
protected:
    mci::ObservableFunctionInterface *_clone() const final { return new MomentsObs(_ndim); }
public:
    explicit MomentsObs(int nd): mci::ObservableFunctionInterface(nd, 2, false) {}
    void observableFunction(const double in[], double out[]) final
    {
        double s = 0., m = 0.;
        for (int i = 0; i < _ndim; ++i) {
            const double a = fabs(in[i]);
            s = s + (a*a)*a;
            m = (a > m) ? a : m;
        }
        out[0] = s;
        out[1] = m;
    }
};

// dependent observables (DependentObservableInterface.hpp:24-54): the log value or the proto values of
// sampling function j as its observation callback last saw them, and the squares of the last values
// of an earlier observable k (read through its accumulator)
class PdfObs final: public mci::ObservableFunctionInterface, public mci::DependentObservableInterface
{
    int _kind, _j;
    const ObservedPDF *_pdf = nullptr;
protected:
    mci::ObservableFunctionInterface *_clone() const final { return new PdfObs(_ndim, _kind, _j); }
public:
    PdfObs(int nd, int kind, int j):
            mci::ObservableFunctionInterface(nd, kind == OBS_PDF_PROTO ? nd : 1, false), mci::DependentObservableInterface(true), _kind(kind), _j(j) {}
    void registerDeps(const mci::SamplingFunctionContainer &pdfcont, const std::vector<mci::AccumulatorInterface *> & /*accuvec*/, int /*selfIdx*/) final
    {
        _pdf = dynamic_cast<const ObservedPDF *>(&pdfcont.getSamplingFunction(_j));
        if (_pdf == nullptr) { throw std::invalid_argument("PdfObs: not a harness sampling function"); }
    }
    void deregisterDeps() final { _pdf = nullptr; }
    void observableFunction(const double /*in*/[], double out[]) final
    {
        if (_pdf == nullptr || (int)_pdf->seen_pv.size() != _pdf->getNProto()) { throw std::runtime_error("PdfObs: no observation yet"); }
        if (_kind == OBS_PDF_PROTO) { for (int i = 0; i < _nobs; ++i) { out[i] = _pdf->seen_pv[i]; } }
        else { out[0] = _pdf->logValue(_pdf->seen_pv.data()); }
    }
};

class DepSquareObs final: public mci::ObservableFunctionInterface, public mci::DependentObservableInterface
{
    int _k;
    const double *_v = nullptr;
protected:
    mci::ObservableFunctionInterface *_clone() const final { return new DepSquareObs(_ndim, _nobs, _k); }
public:
    DepSquareObs(int nd, int nobs, int k): mci::ObservableFunctionInterface(nd, nobs, false), mci::DependentObservableInterface(false), _k(k) {}
    void registerDeps(const mci::SamplingFunctionContainer & /*pdfcont*/, const std::vector<mci::AccumulatorInterface *> &accuvec, int selfIdx) final
    {
        if (_k >= (int)accuvec.size() || !isObsDepValid(accuvec, selfIdx, _k)) { throw std::invalid_argument("DepSquareObs: invalid dependency"); }
        _v = accuvec[_k]->getObsValues();
    }
    void deregisterDeps() final { _v = nullptr; }
    void observableFunction(const double /*in*/[], double out[]) final
    {
        for (int i = 0; i < _nobs; ++i) { out[i] = _v[i] * _v[i]; }
    }
};

std::unique_ptr<mci::SamplingFunctionInterface> makePDF(int nd, int kind, const std::vector<double> &p)
{
    switch (kind) {
    case PDF_GAUSS_ISO: return std::unique_ptr<mci::SamplingFunctionInterface>(new GaussIso(nd));
    case PDF_GAUSS_DIAG: return std::unique_ptr<mci::SamplingFunctionInterface>(new DiagPDF(nd, false, p));
    case PDF_EXP_DIAG: return std::unique_ptr<mci::SamplingFunctionInterface>(new DiagPDF(nd, true, p));
    case PDF_USER_QUARTIC: return std::unique_ptr<mci::SamplingFunctionInterface>(new QuarticPDF(nd, p));
    case PDF_USER_QUARTIC_UPD: return std::unique_ptr<mci::SamplingFunctionInterface>(new QuarticUpdPDF(nd, p));
    case PDF_GMM_ISO: {
        const int K = static_cast<int>(p[0]);
        std::vector<double> w(p.begin() + 1, p.begin() + 1 + K), s(p.begin() + 1 + K, p.begin() + 1 + 2*K);
        std::vector<double> mu(p.begin() + 1 + 2*K, p.end());
        return std::unique_ptr<mci::SamplingFunctionInterface>(new GmmIso(nd, K, w, s, mu));
    }
    case PDF_GMM_FULL: {
        // K, w[K], mu[K][d], P[K][d][d], c[K], L[K][d][d]
        const int K = static_cast<int>(p[0]);
        const size_t dd = (size_t)nd*nd;
        auto at = p.begin() + 1;
        std::vector<double> w(at, at + K); at += K;
        std::vector<double> mu(at, at + (size_t)K*nd); at += (size_t)K*nd;
        std::vector<double> P(at, at + K*dd); at += K*dd;
        std::vector<double> c(at, at + K); at += K;
        std::vector<double> L(at, at + K*dd);
        return std::unique_ptr<mci::SamplingFunctionInterface>(new GmmFull(nd, K, w, mu, P, c, L));
    }
    default: throw std::invalid_argument("unknown pdf kind");
    }
}

mci::EstimatorType estimOf(int e)
{
    switch (e) {
    case 0: return mci::EstimatorType::Noop;
    case 1: return mci::EstimatorType::Uncorrelated;
    case 2: return mci::EstimatorType::Correlated;
    case 3: return mci::EstimatorType::FCBlocker;
    default: return mci::EstimatorType::MJBlocker;
    }
}

std::vector<double> readDoubles(std::istringstream &is, int n)
{
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i) {
        std::string t;
        is >> t;
        v[i] = std::strtod(t.c_str(), nullptr); // accepts decimal and hex-float
    }
    return v;
}

// glibc rand()-driven Metropolis walk with the acceptance rule of a 1-D Gaussian
// orbital product (the kind of data test/ut1/main.cpp feeds to the estimators).
void genWalk(int nmc, int nd, double step, double changeProb, std::vector<double> &x, std::vector<int> &acc,
             std::vector<int> &nch, std::vector<int> &cidx)
{
    auto wf = [nd](const double *p) { double s = 0.; for (int i = 0; i < nd; ++i) { s += p[i]*p[i]; } return exp(-s); };
    x.assign(static_cast<size_t>(nmc)*nd, 0.);
    acc.assign(nmc, 0);
    nch.assign(nmc, 0);
    cidx.assign(static_cast<size_t>(nmc)*nd, 0);
    for (int j = 0; j < nd; ++j) { x[j] = rand()*(1.0/RAND_MAX) - 0.5; }
    acc[0] = 1;
    nch[0] = nd;
    for (int j = 0; j < nd; ++j) { cidx[j] = j; }
    for (int i = 1; i < nmc; ++i) {
        const double *po = &x[static_cast<size_t>(i - 1)*nd];
        double *pn = &x[static_cast<size_t>(i)*nd];
        const double ov = wf(po);
        int n = 0;
        for (int j = 0; j < nd; ++j) {
            bool mv = true;
            if (changeProb < 1.) { mv = (rand()*(1.0/RAND_MAX) <= changeProb); }
            if (mv) { pn[j] = po[j] + 2.*step*(rand()*(1.0/RAND_MAX) - 0.5); cidx[static_cast<size_t>(i)*nd + n] = j; ++n; }
            else { pn[j] = po[j]; }
        }
        nch[i] = n;
        const double nv = wf(pn);
        bool a;
        if (ov == 0) { a = nv != 0; }
        else {
            const double thr = (nv*nv)/(ov*ov);
            a = (thr >= 1.) ? true : (rand()*(1.0/RAND_MAX) <= thr);
        }
        acc[i] = a ? 1 : 0;
        if (!a) { std::copy(po, po + nd, pn); }
    }
}

struct Trace
{
    bool on = false;
    int64_t ncalls = 0;
    std::vector<double> x;
    std::vector<double> rate;
};
} // namespace

int main()
{
    std::unique_ptr<mci::MCI> mci;
    std::unique_ptr<mci::MultiStepMove> msm; // a MultiStepMove being assembled (msmove, mspdf, msapply)
    int ndim = 0;
    std::vector<int> obs_nobs; // components of each observable added, in order
    Trace tr;
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream is(line);
        std::string cmd;
        if (!(is >> cmd) || cmd[0] == '#') { continue; }
        try {
            if (cmd == "create") {
                is >> ndim;
                mci.reset(new mci::MCI(ndim));
                obs_nobs.clear();
            }
            else if (cmd == "seed") { uint64_t s; is >> s; mci->setSeed(s); }
            else if (cmd == "target") { mci->setTargetAcceptanceRate(readDoubles(is, 1)[0]); }
            else if (cmd == "nfind") { int n; is >> n; mci->setNfindMRT2Iterations(n); }
            else if (cmd == "ndecorr") { int64_t n; is >> n; mci->setNdecorrelationSteps(n); }
            else if (cmd == "setx") { auto v = readDoubles(is, ndim); mci->setX(v.data()); }
            else if (cmd == "step") { mci->setMRT2Step(readDoubles(is, 1)[0]); }
            else if (cmd == "stepi") { int i; is >> i; mci->setMRT2Step(i, readDoubles(is, 1)[0]); }
            else if (cmd == "irange") { auto lb = readDoubles(is, ndim); auto ub = readDoubles(is, ndim); mci->setIRange(lb.data(), ub.data()); }
            else if (cmd == "resetdomain") { mci->resetDomain(); }
            else if (cmd == "move") {
                int srrd, veclen, ntypes;
                is >> srrd >> veclen >> ntypes;
                std::vector<int> te(ntypes > 0 ? ntypes : 1);
                for (int i = 0; i < ntypes; ++i) { is >> te[i]; }
                mci->setTrialMove(static_cast<mci::SRRDType>(srrd), veclen, ntypes, ntypes > 1 ? te.data() : nullptr);
            }
            else if (cmd == "movetype") { int t; is >> t; mci->setTrialMove(static_cast<mci::MoveType>(t)); }
            else if (cmd == "umove") {
                int kind, np;
                is >> kind >> np;
                auto p = readDoubles(is, np);
                if (kind == MOVE_USER_PAIR) { mci->setTrialMove(std::unique_ptr<mci::TrialMoveInterface>(new PairMove(ndim, p))); }
                else if (kind == MOVE_USER_SCALE) { mci->setTrialMove(std::unique_ptr<mci::TrialMoveInterface>(new ScaleMove(ndim, p.at(0)))); }
                else { throw std::invalid_argument("unknown user move kind"); }
            }
            else if (cmd == "msmove") {
                // MultiStepMove(ndim, nsteps) around an SRRD move built as MCI::setTrialMove(SRRDType, ...) does
                int nsteps, srrd, veclen, ntypes;
                is >> nsteps >> srrd >> veclen >> ntypes;
                std::vector<int> te(ntypes > 0 ? ntypes : 1);
                for (int i = 0; i < ntypes; ++i) { is >> te[i]; }
                const auto st = static_cast<mci::SRRDType>(srrd);
                int *tep = ntypes > 1 ? te.data() : nullptr;
                auto sub = veclen > 0 ? mci::createSRRDVecMove(st, ndim / veclen, veclen, ntypes, tep) : mci::createSRRDAllMove(st, ndim, ntypes, tep);
                msm.reset(new mci::MultiStepMove(ndim, nsteps));
                msm->setTrialMove(*sub);
            }
            else if (cmd == "mspdf") { int kind, np; is >> kind >> np; auto p = readDoubles(is, np); msm->addSamplingFunction(*makePDF(ndim, kind, p)); }
            else if (cmd == "obsfile") { std::string path; int freq; is >> path >> freq; mci->storeObservablesOnFile(path, freq); }
            else if (cmd == "wlkfile") { std::string path; int freq; is >> path >> freq; mci->storeWalkerPositionsOnFile(path, freq); }
            else if (cmd == "msapply") { mci->setTrialMove(std::unique_ptr<mci::TrialMoveInterface>(msm.release())); }
            else if (cmd == "pdf") { int kind, np; is >> kind >> np; auto p = readDoubles(is, np); mci->addSamplingFunction(makePDF(ndim, kind, p)); }
            else if (cmd == "poppdf") { mci->popSamplingFunction(); }
            else if (cmd == "obs") {
                int kind, np, bs, nskip, eq, est;
                is >> kind >> np;
                auto p = readDoubles(is, np);
                is >> bs >> nskip >> eq >> est;
                std::unique_ptr<mci::ObservableFunctionInterface> o;
                if (kind == OBS_PDF_LOG || kind == OBS_PDF_PROTO) { o.reset(new PdfObs(ndim, kind, p.empty() ? 0 : static_cast<int>(p[0]))); }
                else if (kind == OBS_DEP_SQUARE) {
                    const int k = static_cast<int>(p.at(0));
                    o.reset(new DepSquareObs(ndim, obs_nobs.at(k), k));
                }
                else if (kind == OBS_USER_MOMENTS) { o.reset(new MomentsObs(ndim)); }
                else { o.reset(new CatObs(ndim, kind, p)); }
                obs_nobs.push_back(o->getNObs());
                mci->addObservable(std::move(o), bs, nskip, eq != 0, estimOf(est));
            }
            else if (cmd == "popobs") { mci->popObservable(); obs_nobs.pop_back(); }
            else if (cmd == "clearobs") { mci->clearObservables(); obs_nobs.clear(); }
            else if (cmd == "movex") { mci->moveX(); }
            else if (cmd == "newrandomx") { mci->newRandomX(); }
            else if (cmd == "centerx") { mci->centerX(); }
            else if (cmd == "trace") {
                // record x and the running acceptance rate at every callback (MCIntegrator.cpp:265,347)
                tr = Trace();
                tr.on = true;
                mci->setCallback([&tr](const mci::MCI &m) {
                    ++tr.ncalls;
                    for (int i = 0; i < m.getNDim(); ++i) { tr.x.push_back(m.getX(i)); }
                    tr.rate.push_back(m.getAcceptanceRate());
                });
            }
            else if (cmd == "untrace") { mci->clearCallback(); tr.on = false; }
            else if (cmd == "integrate") {
                int64_t nmc; int df, dd;
                is >> nmc >> df >> dd;
                const int nd = mci->getNObsDim();
                std::vector<double> avg(nd > 0 ? nd : 1), err(nd > 0 ? nd : 1);
                tr.ncalls = 0; tr.x.clear(); tr.rate.clear();
                auto t0 = std::chrono::steady_clock::now();
                mci->integrate(nmc, avg.data(), err.data(), df != 0, dd != 0);
                auto t1 = std::chrono::steady_clock::now();
                std::vector<double> steps(mci->getTrialMove().getNStepSizes());
                for (size_t i = 0; i < steps.size(); ++i) { steps[i] = mci->getMRT2Step(static_cast<int>(i)); }
                std::cout << "{\"cmd\":\"integrate\",\"avg\":" << hxv(avg.data(), nd) << ",\"err\":" << hxv(err.data(), nd)
                          << ",\"rate\":" << hx(mci->getAcceptanceRate()) << ",\"steps\":" << hxv(steps.data(), static_cast<int>(steps.size()))
                          << ",\"x\":" << hxv(mci->getX(), ndim)
                          << ",\"seconds\":" << std::chrono::duration<double>(t1 - t0).count();
                if (tr.on) {
                    std::cout << ",\"trace_x\":" << hxv(tr.x.data(), static_cast<int>(tr.x.size()))
                              << ",\"trace_rate\":" << hxv(tr.rate.data(), static_cast<int>(tr.rate.size()));
                }
                std::cout << "}" << std::endl;
            }
            else if (cmd == "getx") { std::cout << "{\"cmd\":\"getx\",\"x\":" << hxv(mci->getX(), ndim) << "}" << std::endl; }
            else if (cmd == "estimate") {
                // estimate EST N ND v0 v1 ... (N*ND values)
                int est; int64_t n; int nd;
                is >> est >> n >> nd;
                auto x = readDoubles(is, static_cast<int>(n*nd));
                std::vector<double> a(nd), e(nd);
                auto fn = mci::createEstimator(estimOf(est));
                fn(n, nd, x.data(), a.data(), e.data());
                std::cout << "{\"cmd\":\"estimate\",\"avg\":" << hxv(a.data(), nd) << ",\"err\":" << hxv(e.data(), nd) << "}" << std::endl;
            }
            else if (cmd == "walkdata") {
                // walkdata SEED NMC ND STEP CHANGEPROB : emit a rand()-driven walk and the reference's
                // estimator and accumulator outputs on it (the checks of test/ut1/main.cpp)
                unsigned seed; int nmc, nd; double step, cp;
                is >> seed >> nmc >> nd >> step >> cp;
                srand(seed);
                std::vector<double> x; std::vector<int> acc, nch, cidx;
                genWalk(nmc, nd, step, cp, x, acc, nch, cidx);
                std::cout << "{\"cmd\":\"walkdata\",\"x\":" << hxv(x.data(), static_cast<int>(x.size())) << ",\"acc\":[";
                for (int i = 0; i < nmc; ++i) { std::cout << (i ? "," : "") << acc[i]; }
                std::cout << "],\"nch\":[";
                for (int i = 0; i < nmc; ++i) { std::cout << (i ? "," : "") << nch[i]; }
                std::cout << "],\"cidx\":[";
                for (size_t i = 0; i < cidx.size(); ++i) { std::cout << (i ? "," : "") << cidx[i]; }
                std::cout << "]";
                std::vector<double> a(nd), e(nd);
                const char *names[] = {"noop", "uncorr", "corr", "fc", "mj"};
                for (int est = 1; est <= 4; ++est) {
                    mci::createEstimator(estimOf(est))(nmc, nd, x.data(), a.data(), e.data());
                    std::cout << ",\"" << names[est] << "_nd\":{\"avg\":" << hxv(a.data(), nd) << ",\"err\":" << hxv(e.data(), nd) << "}";
                }
                std::vector<double> col(nmc);
                for (int j = 0; j < nd; ++j) {
                    for (int i = 0; i < nmc; ++i) { col[i] = x[static_cast<size_t>(i)*nd + j]; }
                    double a1, e1;
                    mci::OneDimUncorrelatedEstimator(nmc, col.data(), a1, e1);
                    std::cout << ",\"uncorr_1d_" << j << "\":[" << hx(a1) << "," << hx(e1) << "]";
                    mci::OneDimBlockEstimator(nmc, col.data(), 2048, a1, e1);
                    std::cout << ",\"block_1d_" << j << "\":[" << hx(a1) << "," << hx(e1) << "]";
                    mci::OneDimFCBlockerEstimator(nmc, col.data(), a1, e1);
                    std::cout << ",\"fc_1d_" << j << "\":[" << hx(a1) << "," << hx(e1) << "]";
                    mci::MJBlockerEstimator(nmc, 1, col.data(), &a1, &e1);
                    std::cout << ",\"mj_1d_" << j << "\":[" << hx(a1) << "," << hx(e1) << "]";
                }
                mci::MultiDimBlockEstimator(nmc, nd, x.data(), 2048, a.data(), e.data());
                std::cout << ",\"block_nd\":{\"avg\":" << hxv(a.data(), nd) << ",\"err\":" << hxv(e.data(), nd) << "}";
                // accumulators fed with the walk (test/ut1/main.cpp:108-128 feeding scheme)
                CatObs xobs(nd, OBS_X, {});
                CatObs x2obs(nd, OBS_X2, {});
                struct A { std::string name; std::unique_ptr<mci::AccumulatorInterface> acc; };
                std::vector<A> accs;
                accs.push_back({"simple1", std::unique_ptr<mci::AccumulatorInterface>(new mci::SimpleAccumulator(xobs, 1))});
                accs.push_back({"simple2", std::unique_ptr<mci::AccumulatorInterface>(new mci::SimpleAccumulator(xobs, 2))});
                accs.push_back({"block16", std::unique_ptr<mci::AccumulatorInterface>(new mci::BlockAccumulator(xobs, 1, 16))});
                accs.push_back({"block8s2", std::unique_ptr<mci::AccumulatorInterface>(new mci::BlockAccumulator(xobs, 2, 8))});
                accs.push_back({"full1", std::unique_ptr<mci::AccumulatorInterface>(new mci::FullAccumulator(xobs, 1))});
                accs.push_back({"full2", std::unique_ptr<mci::AccumulatorInterface>(new mci::FullAccumulator(xobs, 2))});
                accs.push_back({"x2full3", std::unique_ptr<mci::AccumulatorInterface>(new mci::FullAccumulator(x2obs, 3))});
                accs.push_back({"x2block4s3", std::unique_ptr<mci::AccumulatorInterface>(new mci::BlockAccumulator(x2obs, 3, 4))});
                for (auto &ac : accs) {
                    const int64_t steps = (ac.name == "x2block4s3") ? static_cast<int64_t>(nmc) - 4 : nmc; // NAccu divisible by 4
                    ac.acc->allocate(steps);
                    mci::WalkerState wlk(nd, true);
                    for (int64_t i = 0; i < steps; ++i) {
                        std::copy(&x[static_cast<size_t>(i)*nd], &x[static_cast<size_t>(i + 1)*nd], wlk.xnew);
                        wlk.nchanged = nch[i];
                        std::copy(&cidx[static_cast<size_t>(i)*nd], &cidx[static_cast<size_t>(i + 1)*nd], wlk.changedIdx);
                        wlk.accepted = acc[i] != 0;
                        ac.acc->accumulate(wlk);
                    }
                    ac.acc->finalize();
                    std::cout << ",\"accu_" << ac.name << "\":" << hxv(ac.acc->getData(), static_cast<int>(ac.acc->getNData()));
                }
                std::cout << "}" << std::endl;
            }
            else if (cmd == "bench") {
                // bench W NMC : W independent walkers (one MCI each, seeds seed0+w), full integrate each,
                // timed; prints walker-steps/s of the sampling phase and the total
                int64_t W, nmc; uint64_t seed0;
                is >> W >> nmc >> seed0;
                // the configuration lines after "bench" are replayed for every walker
                std::vector<std::string> setup;
                std::string l2;
                while (std::getline(std::cin, l2) && l2 != "endbench") { setup.push_back(l2); }
                double tot = 0., avgsum = 0.;
                int64_t steps_total = 0;
                std::vector<double> asum, a2sum, e2sum;
                for (int64_t w = 0; w < W; ++w) {
                    std::unique_ptr<mci::MCI> m;
                    std::unique_ptr<GmmFull> start; // exact start when the target is GMM_FULL
                    int nd0 = 0;
                    for (auto &sl : setup) {
                        std::istringstream ss(sl);
                        std::string c2; ss >> c2;
                        if (c2 == "create") { ss >> nd0; m.reset(new mci::MCI(nd0)); m->setSeed(seed0 + static_cast<uint64_t>(w)); }
                        else if (c2 == "pdf") {
                            int kind, np; ss >> kind >> np; auto p = readDoubles(ss, np);
                            auto f = makePDF(nd0, kind, p);
                            if (kind == PDF_GMM_FULL) { start.reset(static_cast<GmmFull *>(f->clone().release())); }
                            m->addSamplingFunction(std::move(f));
                        }
                        else if (c2 == "obs") {
                            int kind, np, bs, nskip, eq, est; ss >> kind >> np; auto p = readDoubles(ss, np); ss >> bs >> nskip >> eq >> est;
                            m->addObservable(std::unique_ptr<mci::ObservableFunctionInterface>(new CatObs(nd0, kind, p)), bs, nskip, eq != 0, estimOf(est));
                        }
                        else if (c2 == "move") { int srrd, vl, nt; ss >> srrd >> vl >> nt; m->setTrialMove(static_cast<mci::SRRDType>(srrd), vl, nt, nullptr); }
                        else if (c2 == "nfind") { int n; ss >> n; m->setNfindMRT2Iterations(n); }
                        else if (c2 == "ndecorr") { int64_t n; ss >> n; m->setNdecorrelationSteps(n); }
                    }
                    std::vector<double> a(m->getNObsDim()), e(m->getNObsDim());
                    if (start) {
                        std::mt19937_64 g(seed0 + static_cast<uint64_t>(w) + 0x5eed0000ull);
                        std::vector<double> x0(nd0);
                        start->draw(g, x0.data());
                        m->setX(x0.data());
                    }
                    auto t0 = std::chrono::steady_clock::now();
                    m->integrate(nmc, a.data(), e.data(), true, true);
                    auto t1 = std::chrono::steady_clock::now();
                    tot += std::chrono::duration<double>(t1 - t0).count();
                    steps_total += nmc;
                    avgsum += a.empty() ? 0. : a[0];
                    asum.resize(a.size(), 0.);
                    a2sum.resize(a.size(), 0.);
                    e2sum.resize(e.size(), 0.);
                    for (size_t i = 0; i < a.size(); ++i) { asum[i] += a[i]; a2sum[i] += a[i]*a[i]; e2sum[i] += e[i]*e[i]; }
                }
                std::cout << "{\"cmd\":\"bench\",\"walkers\":" << W << ",\"seconds\":" << tot << ",\"walker_steps\":" << steps_total
                          << ",\"avg0_mean\":" << avgsum/W << ",\"avg_sum\":" << hxv(asum.data(), static_cast<int>(asum.size()))
                          << ",\"avg2_sum\":" << hxv(a2sum.data(), static_cast<int>(a2sum.size()))
                          << ",\"err2_sum\":" << hxv(e2sum.data(), static_cast<int>(e2sum.size())) << "}" << std::endl;
            }
            else { std::cout << "{\"cmd\":\"error\",\"what\":\"unknown command " << cmd << "\"}" << std::endl; }
        }
        catch (std::invalid_argument &ex) { std::cout << "{\"cmd\":\"exception\",\"type\":\"invalid_argument\",\"what\":\"" << ex.what() << "\"}" << std::endl; }
        catch (std::domain_error &ex) { std::cout << "{\"cmd\":\"exception\",\"type\":\"domain_error\",\"what\":\"" << ex.what() << "\"}" << std::endl; }
        catch (std::runtime_error &ex) { std::cout << "{\"cmd\":\"exception\",\"type\":\"runtime_error\",\"what\":\"" << ex.what() << "\"}" << std::endl; }
    }
    return 0;
}
