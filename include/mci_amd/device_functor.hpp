/*
 * device_functor.hpp — what the source of a caller's device functor sees.
 *
 * The reference's user plugins are C++ subclasses whose virtual methods take the walker as plain
 * arrays (SamplingFunctionInterface.hpp:74-97, ProtoFunctionInterface.hpp:24-65,
 * ObservableFunctionInterface.hpp:48-61).  Here the walkers are structure-of-arrays in device
 * memory, one walker per lane, so a functor receives each vector as an MciVec: element i of this
 * walker's vector, read or written as v[i].  A functor source (HIP C++) is handed to
 * mci_add_sampling_function_device / mci_add_observable_device with a name NAME and defines
 *
 *   __device__ void   NAME_proto(int ndim, MciVec x, const double *par, MciVec proto);
 *   __device__ double NAME_acceptance(int nproto, MciVec protoold, MciVec protonew, const double *par);
 *   __device__ double NAME_sampling(int nproto, MciVec proto, const double *par);
 *
 * for a sampling function (protoFunction, acceptanceFunction, samplingFunction) and, if it
 * overrides updatedAcceptance (SamplingFunctionInterface.hpp:74-91: the acceptance after a move of
 * some coordinates, updating only the proto values those change),
 *
 *   __device__ double NAME_updated_acceptance(int ndim, int nproto, MciVec xnew, MciChanged changed,
 *                                             MciVec protoold, MciVec protonew, const double *par);
 *
 * (without it the proto values are recomputed in full, the interface's default), or
 *
 *   __device__ void   NAME_observable(int ndim, MciVec x, const double *par, MciVec out);
 *
 * for an observable (observableFunction), or
 *
 *   __device__ double NAME_move(int ndim, MciVec xnew, MciMoved &moved, MciVec protoold,
 *                               MciVec protonew, MciVec steps, MciRng &rng, const double *par);
 *   __device__ void   NAME_proto(int ndim, MciVec x, const double *par, MciVec proto);  (with proto values)
 *
 * for a trial move (TrialMoveInterface.hpp:16-70: trialMove and protoFunction; xnew holds the
 * current position on entry, the move changes the range moved.first .. moved.first + moved.n - 1
 * (all coordinates unless it says otherwise) and returns its acceptance factor; steps[i] are its
 * step sizes, draws come from mci_uniform(rng)), par being the parameters given with it.  The library
 * compiles the sources together with its generic kernels for the context's GPU when the chain
 * first needs them.  This header is included ahead of the sources; it needs nothing else.
 */
#pragma once

#include <stdint.h>

struct MciVec
{
    double *p;  // element 0 of this walker's vector
    int64_t s;  // distance between consecutive elements (the walker count)
    __device__ double &operator[](int i) const { return p[(int64_t)i * s]; }
};

// the coordinates a move changed, in ascending order (WalkerState::changedIdx / nchanged):
// changed[i] for i < changed.n
struct MciChanged
{
    int first, n;
    __device__ int operator[](int i) const { return first + i; }
};

// what a trial move changed: coordinates first .. first + n - 1 (WalkerState::nchanged / changedIdx
// for one ascending range; unchanged coordinates inside it are harmless)
struct MciMoved
{
    int first, n;
};

// the walker's random stream; mci_uniform(rng) is a uniform on [0, 1), the draw of
// std::uniform_real_distribution<double>(0, 1) on MCI's generator in the reference-stream mode
struct MciRng
{
    const void *c;
    void *L;
};
__device__ double mci_uniform(MciRng &rng);
