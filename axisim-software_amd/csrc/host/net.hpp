// net.hpp -- the learned initial guess for the pressure-correction solve: a
// message-passing network on a cell graph, trained on the GPU by its own
// forward and backward kernels (mpnn.hip).  Build extension with no reference
// counterpart.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../hip/mpnn.hpp"
#include "solver.hpp"

namespace axs {

// Host copy of a graph's CSR plus the reverse-slot map the transposed
// aggregation needs (rev[k] = slot of cell j in the row of nb(k)).
std::vector<int> reverseSlots(int n, const std::vector<int>& start, const std::vector<int>& nb);

// A graph resident on the device with its weight arrays.
struct NetGraphDev {
    NetGraph g;
    Arena arena;
    void build(int n, const std::vector<int>& start, const std::vector<int>& nb);
};

// Activations and gradients for graphs of up to `rows` padded rows.
struct NetWork {
    int rows = 0, layers = 0;
    Arena arena;
    float *r = nullptr, *y = nullptr, *dy = nullptr, *dH = nullptr, *dA = nullptr, *partial = nullptr;
    uint16_t *H0 = nullptr, *dZ = nullptr;
    std::vector<uint16_t*> A, H;  // per layer: [H_{l-1} | M_{l-1}] and H_l
    double *sigma = nullptr, *dpartial = nullptr, *loss = nullptr;
    double *x = nullptr, *Ax = nullptr, *scal = nullptr;  // guess, A guess, [x.b, x.Ax, alpha]
    void ensure(int nPad, int layers);
};

class Mpnn {
public:
    Mpnn(int layers, uint64_t seed, int device);
    ~Mpnn();
    int layers() const { return layers_; }
    int numParams() const { return numParams_; }
    void getParams(float* out) const;
    void setParams(const float* in);

    // training pairs: the system (CSR, AC, AF, b) and the target solution x*
    int addSample(int n, const int* start, const int* nb, const double* AC, const double* AF, const double* b,
                  const double* target);
    void clearSamples();
    int samples() const { return (int)samples_.size(); }
    // loss of one sample and its gradient (FP32, parameter order)
    double lossGrad(int sample, float* grad);
    // Adam over the samples in shuffled order, `batch` samples per step
    void train(int epochs, int batch, double lr, uint64_t seed, double* lossPerEpoch);
    // the guess for one system (host arrays)
    void predict(int n, const int* start, const int* nb, const double* AC, const double* AF, const double* b,
                 double* x);
    // the guess for a system resident on the device, enqueued on s (no
    // allocation: the graph and work buffers are set up beforehand)
    void enqueueForward(hipStream_t s, const NetGraph& g, const double* AC, const double* AF, const double* b,
                        double* x, NetWork& w);

private:
    struct Sample {
        NetGraphDev graph;
        Arena arena;
        double *AC = nullptr, *AF = nullptr, *b = nullptr;
        double estar = 0.0;  // b.x* = |x*|_A^2, the energy of the solution
    };
    int K(int l) const { return l == 0 ? 2 * kNetIn : 2 * kNetHidden; }
    size_t offW(int l) const;
    size_t offOut() const { return offW(layers_); }
    void castWeights(hipStream_t s);
    void forward(hipStream_t s, const NetGraph& g, const double* AC, const double* b, NetWork& w, bool scale);
    void backward(hipStream_t s, const NetGraph& g, const Sample& smp, NetWork& w);

    int layers_ = 4, numParams_ = 0, device_ = 0, step_ = 0;
    hipStream_t stream_ = nullptr;
    Arena arena_;
    float *P_ = nullptr, *G_ = nullptr, *M_ = nullptr, *V_ = nullptr;
    std::vector<uint16_t*> Wb_, WTb_;
    std::vector<Sample*> samples_;
    NetWork work_;
};

}  // namespace axs
