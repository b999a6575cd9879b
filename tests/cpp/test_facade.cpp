// The checks of the reference's unit tests test/ut2, ut3 and ut4 (main.cpp of each), run
// through the C++ facade include/mci_amd/MCIntegrator.hpp on the GPU library.  With the
// default random mode (per-walker std::mt19937_64 streams) the chains are the reference's
// own, so the seeded assertions hold exactly as they do there.  Also checks the facade's
// exception mapping and a many-walker context.
//
// Built by tests/cpp/Makefile; run by tests/test_cpp_facade.py.  Prints one line per section
// and exits non-zero on the first failed check.
#include "mci_amd/MCIntegrator.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

using namespace mci;

static int g_checks = 0;
#define CHECK(cond)                                                                          \
    do {                                                                                     \
        ++g_checks;                                                                          \
        if (!(cond)) {                                                                       \
            std::fprintf(stderr, "CHECK failed at %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)

template <class E, class F>
static bool throws(F f)
{
    try {
        f();
    }
    catch (const E &) {
        return true;
    }
    catch (...) {
        return false;
    }
    return false;
}

static void ut2()
{
    MCI mci(3);
    for (int i = 0; i < 3; ++i) { CHECK(mci.getX(i) == 0.); }
    mci.moveX();
    bool same = true;
    for (int i = 0; i < 3; ++i) { same = same && mci.getX(i) == 0.; }
    CHECK(!same);

    mci.setIRange(-5., 5.); // periodic box, as OrthoPeriodicDomain(3, -5., 5.)
    mci.centerX();
    for (int i = 0; i < 3; ++i) { CHECK(mci.getX(i) == 0.); }
    mci.newRandomX();
    for (int i = 0; i < 3; ++i) { CHECK(mci.getX(i) != 0.); }
    mci.resetDomain();

    const int NMC = 10000;
    const double CORRECT = 0.5;
    mci.addSamplingFunction(ThreeDimGaussianPDF());
    mci.addObservable(XSquared());
    double x[3] = {5., -5., 10.};
    mci.setSeed(5649871);
    double average, error;

    mci.setX(x);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) > 2. * error);

    mci.setX(x);
    mci.integrate(NMC, &average, &error);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.setIRange(-5., 5.);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.popSamplingFunction();
    mci.popObservable();
    mci.addObservable(GaussXSquared(), 1, 1, false, false);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);
    std::printf("ut2 ok: GaussXSquared without sampling function %.6f +- %.6f\n", average, error);
}

static void ut3()
{
    const int NMC = 10000;
    const double CORRECT = 0.5;
    MCI mci(3);
    mci.setSeed(5649871);
    mci.addSamplingFunction(ThreeDimGaussianPDF());
    mci.addObservable(XSquared());
    mci.addObservable(XYZSquared());
    double x[3] = {5., -5., 10.};
    double average[4], error[4];

    mci.setX(x);
    mci.integrate(NMC, average, error, false, false);
    for (int i = 0; i < mci.getNObsDim(); ++i) { CHECK(std::fabs(average[i] - CORRECT) > 2. * error[i]); }

    mci.setX(x);
    mci.integrate(NMC, average, error);
    for (int i = 0; i < mci.getNObsDim(); ++i) { CHECK(std::fabs(average[i] - CORRECT) < 2. * error[i]); }

    mci.integrate(NMC, average, error, false, false);
    for (int i = 0; i < mci.getNObsDim(); ++i) { CHECK(std::fabs(average[i] - CORRECT) < 2. * error[i]); }
    std::printf("ut3 ok: nobsdim %d\n", mci.getNObsDim());
}

static void ut4()
{
    const int NMC = 16384;
    const double CORRECT = 0.5;
    MCI mci(3);
    mci.setSeed(1337);
    mci.addSamplingFunction(ThreeDimGaussianPDF());
    mci.addObservable(XSquared());
    double x[3] = {5., -5., 10.};
    double average, error;

    mci.setNfindMRT2Iterations(20);
    mci.setNdecorrelationSteps(2000);
    CHECK(mci.getNfindMRT2Iterations() == 20 && mci.getNdecorrelationSteps() == 2000);

    mci.setX(x);
    mci.integrate(NMC, &average, &error, true, true);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.clearObservables();
    mci.addObservable(XSquared(), 16);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.clearObservables();
    mci.addObservable(XSquared(), 8, 2);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.clearObservables();
    mci.addObservable(XSquared(), 8, 2, false, true);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);

    mci.clearObservables();
    mci.addObservable(XSquared(), 8, 2, false, EstimatorType::Correlated);
    mci.integrate(NMC, &average, &error, false, false);
    CHECK(std::fabs(average - CORRECT) < 2. * error);
    std::printf("ut4 ok: acceptance %.4f\n", mci.getAcceptanceRate());
}

static void exceptions()
{
    // the reference's exception types for misuse (MCIntegrator.cpp:46,457-460, BlockAccumulator.cpp:12, MJBlocker.cpp:29, Factories.hpp:65)
    CHECK(throws<std::invalid_argument>([] { MCI m(0); }));
    MCI mci(3);
    CHECK(throws<std::invalid_argument>([&] { mci.addObservable(X2(2)); }));
    CHECK(throws<std::invalid_argument>([&] { mci.addObservable(XSquared(), 0, 1, true, EstimatorType::Noop); }));
    CHECK(throws<std::invalid_argument>([] { selectEstimatorType(true, false); }));
    mci.addObservable(XSquared());
    double a, e;
    CHECK(throws<std::domain_error>([&] { mci.integrate(1000, &a, &e); })); // infinite domain, no sampling function
    mci.clearObservables();
    mci.addSamplingFunction(ThreeDimGaussianPDF());
    mci.addObservable(XSquared(), 16);
    CHECK(throws<std::invalid_argument>([&] { mci.integrate(1000, &a, &e); })); // 1000 % 16 != 0
    CHECK(throws<std::invalid_argument>([] {
        std::vector<double> x(3, 1.);
        double av[3], er[3];
        MJBlockerEstimator(3, 1, x.data(), av, er); // MJBlocker needs a power of 2
    }));
    std::printf("exceptions ok\n");
}

static void walkers()
{
    // 256 independent chains on one context, combined as the reference's MPIMCI combines ranks
    const int64_t W = 256;
    MCI mci(3, W);
    mci.setSeed(2024);
    mci.setRandomMode(RandomMode::Philox);
    mci.addSamplingFunction(Gauss(3));
    mci.addObservable(X2(3), 64, 1, true, EstimatorType::MJBlocker);
    double average[3], error[3];
    mci.integrate(64 * 128, average, error);
    std::vector<double> wa, we;
    mci.getWalkerResults(wa, we);
    CHECK((int64_t)wa.size() == 3 * W);
    for (int i = 0; i < 3; ++i) {
        double s = 0., s2 = 0.;
        for (int64_t w = 0; w < W; ++w) {
            s += wa[w * 3 + i];
            s2 += we[w * 3 + i] * we[w * 3 + i];
        }
        CHECK(std::fabs(average[i] - s / W) < 1e-12);
        CHECK(std::fabs(error[i] - std::sqrt(s2) / W) < 1e-12 * error[i] + 1e-300);
        CHECK(std::fabs(average[i] - 0.5) < 4. * error[i]); // <x_i^2> = 1/2 under exp(-sum x^2)
    }
    std::printf("walkers ok: <x0^2> = %.6f +- %.6f over %lld walkers\n", average[0], error[0], (long long)W);
}

static void multistep()
{
    // a MultiStepMove of 4 Gaussian single-coordinate steps under a cheap own Gaussian, around
    // exp(-sum x^2): unbiased, and the change rate is the contained move's times the steps
    const int64_t W = 128;
    MCI mci(3, W);
    mci.setSeed(77);
    mci.addSamplingFunction(Gauss(3));
    MultiStepMove ms(3, 4);
    ms.setTrialMove(SRRDType::Gaussian, 1);
    ms.addSamplingFunction(AnisotropicGaussPDF({1.5, 1.5, 1.5}));
    mci.setTrialMove(ms);
    double rate = 0.;
    check(mci_get_change_rate(mci.handle(), &rate));
    CHECK(std::fabs(rate - 1.) < 1e-15); // min(1, 4 / 3)
    mci.addObservable(X2(3), 16, 1, false, EstimatorType::Uncorrelated);
    double average[3], error[3];
    mci.integrate(16 * 128, average, error);
    for (int i = 0; i < 3; ++i) { CHECK(std::fabs(average[i] - 0.5) < 4. * error[i]); }
    bool threw = false;
    try {
        MultiStepMove bad(3, 2);
        bad.addSamplingFunction(Gauss(2));
    }
    catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw);
    std::printf("multistep ok: <x0^2> = %.6f +- %.6f\n", average[0], error[0]);
}

static void callback()
{
    // the callback sees the start of the sample() call and every step: n + 1 calls without
    // calibration and decorrelation, rate 0 at the start, positions as getX returns them
    MCI mci(2);
    mci.setSeed(9);
    mci.addSamplingFunction(Gauss(2));
    mci.addObservable(X2(2), 1, 1, false, EstimatorType::Uncorrelated);
    std::vector<double> rates, xs;
    mci.setCallback([&](const MCI &m) {
        rates.push_back(m.getAcceptanceRate());
        xs.push_back(m.getX(0));
    });
    double average[2], error[2];
    mci.integrate(100, average, error, false, false);
    CHECK(rates.size() == 101);
    CHECK(rates[0] == 0.);
    bool in01 = true;
    for (double r : rates) { in01 = in01 && r >= 0. && r <= 1.; }
    CHECK(in01);
    mci.clearCallback();
    mci.integrate(100, average, error, false, false);
    CHECK(rates.size() == 101);
    bool threw = false;
    mci.setCallback([](const MCI &) { throw std::runtime_error("from the callback"); });
    try {
        mci.integrate(10, average, error, false, false);
    }
    catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    std::printf("callback ok: %zu calls\n", rates.size());
}

static void learned_mixture_move()
{
    // a bimodal isotropic mixture (weights 0.3 / 0.7 at x_0 = -8 / +8) whose modes plain moves do not
    // cross; every walker starts in the light mode; the mixed move (an independence draw from a
    // mixture with the weights reversed, with probability 0.25) gives P(x_0 > 0) = 0.7
    const int64_t W = 256;
    MCI mci(2, W);
    mci.setSeed(99);
    mci.setRandomMode(RandomMode::Philox);
    mci.addSamplingFunction(GaussianMixturePDF(2, {0.3, 0.7}, {0.8, 1.3}, {-8., 0., 8., 0.}));
    LearnedMixtureMove lm(2, 0.25);
    lm.addComponent(0.75, {-7.7, 0.2}, {1.1, 1.0});
    lm.addComponent(0.25, {8.3, -0.1}, {1.4, 1.5});
    lm.setTrialMove(SRRDType::Uniform, 1);
    mci.setTrialMove(lm);
    std::vector<double> x0(2 * W, 0.);
    for (int64_t w = 0; w < W; ++w) { x0[2 * w] = -8.; }
    check(mci_set_x(mci.handle(), x0.data(), (int64_t)x0.size()));
    mci.addObservable(Indicator(2, 0), 64, 1, false, EstimatorType::Uncorrelated);
    double average = 0., error = 0.;
    mci.integrate(64 * 64, &average, &error);
    CHECK(std::fabs(average - 0.7) < 5. * error + 0.02);
    CHECK(throws<std::invalid_argument>([] { LearnedMixtureMove bad(2, 0.5); bad.addComponent(1., {0.}, {1., 1.}); }));
    std::printf("learned mixture move ok: P(x0 > 0) = %.4f +- %.4f\n", average, error);
}

// a reference user's own plugins written as device code: exp(-|x|^2) in ThreeDimGaussianPDF's
// arithmetic and x_i^2 in XYZSquared's give the catalogue's chain bit for bit
static void device_functors()
{
    const char *pdf = R"(
__device__ void gauss3_proto(int ndim, MciVec x, const double *par, MciVec proto)
{
    double s = x[0] * x[0];
    for (int i = 1; i < ndim; ++i) { s = s + x[i] * x[i]; }
    proto[0] = s;
}
__device__ double gauss3_acceptance(int nproto, MciVec po, MciVec pn, const double *par) { return exp(-pn[0] + po[0]); }
__device__ double gauss3_sampling(int nproto, MciVec pv, const double *par) { return exp(-pv[0]); }
)";
    const char *obs = R"(
__device__ void sq_observable(int ndim, MciVec x, const double *par, MciVec out)
{
    for (int i = 0; i < ndim; ++i) { out[i] = x[i] * x[i]; }
}
)";
    double a1[3], e1[3], a2[3], e2[3];
    {
        MCI mci(3);
        mci.setSeed(5649871);
        mci.addSamplingFunction(DeviceSamplingFunction("gauss3", pdf, 3, 1));
        mci.addObservable(DeviceObservable("sq", obs, 3, 3), 16, 1, true, EstimatorType::MJBlocker);
        mci.integrate(16 * 1024, a1, e1);
    }
    {
        MCI mci(3);
        mci.setSeed(5649871);
        mci.addSamplingFunction(ThreeDimGaussianPDF());
        mci.addObservable(XYZSquared(), 16, 1, true, EstimatorType::MJBlocker);
        mci.integrate(16 * 1024, a2, e2);
    }
    for (int i = 0; i < 3; ++i) {
        CHECK(a1[i] == a2[i] && e1[i] == e2[i]);
        CHECK(std::fabs(a1[i] - 0.5) < 4. * e1[i]);
    }
    MCI bad(2);
    bad.addSamplingFunction(DeviceSamplingFunction("nope", "__device__ void nope_proto(int n, MciVec x, const double *p, MciVec q) { q[0] = z; }", 2, 1));
    bad.addObservable(DeviceObservable("sq", obs, 2, 2), 16, 1, false, EstimatorType::Uncorrelated);
    double a[2], e[2];
    CHECK(throws<std::invalid_argument>([&] { bad.integrate(256, a, e); }));
    std::printf("device functors ok: <x_i^2> = %.4f %.4f %.4f\n", a1[0], a1[1], a1[2]);

    // a trial move of the caller's (two adjacent coordinates, two step types) on exp(-|x|^2): <x_i^2> = 1/2
    DeviceTrialMove pm;
    pm.name = "pair";
    pm.source = R"(
__device__ double pair_move(int ndim, MciVec x, MciMoved &moved, MciVec po, MciVec pn, MciVec steps, MciRng &rng, const double *par)
{
    int i = (int)(mci_uniform(rng) * ndim);
    if (i > ndim - 1) { i = ndim - 1; }
    const int n = (i + 1 < ndim) ? 2 : 1;
    for (int k = 0; k < n; ++k) { const int j = i + k; x[j] = x[j] + steps[j < ndim / 2 ? 0 : 1] * (2.0 * mci_uniform(rng) - 1.0); }
    moved.first = i;
    moved.n = n;
    return 1.0;
}
)";
    pm.steps = {0.3, 0.6};
    pm.typeEnds = {2, 4};
    pm.changeRate = 0.5;
    MCI mm(4, 256);
    mm.setSeed(11);
    mm.setTrialMove(pm);
    mm.addSamplingFunction(Gauss(4));
    mm.addObservable(X2(4), 32, 1, true, EstimatorType::MJBlocker);
    double a4[4], e4[4];
    mm.integrate(32 * 64, a4, e4);
    for (int i = 0; i < 4; ++i) { CHECK(std::fabs(a4[i] - 0.5) < 4. * e4[i]); }
    std::printf("device trial move ok: <x_i^2> = %.4f %.4f %.4f %.4f\n", a4[0], a4[1], a4[2], a4[3]);
}

int main()
{
    int ndev = 0;
    check(mci_device_count(&ndev));
    if (ndev < 1) {
        std::fprintf(stderr, "no GPU visible\n");
        return 2;
    }
    std::printf("%s\n", mci_version());
    std::fflush(stdout);
    ut2();
    ut3();
    ut4();
    exceptions();
    walkers();
    multistep();
    callback();
    learned_mixture_move();
    device_functors();
    std::printf("all %d checks passed\n", g_checks);
    return 0;
}
