// net.cpp -- host side of the learned initial guess (net.hpp).
#include "net.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace axs {

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double normal(uint64_t& s) {  // Box-Muller
    const double u = uniform01(s), v = uniform01(s);
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
}

void checkCSR(int n, const int* start, const int* nb) {
    if (n <= 0 || !start || !nb) throw std::invalid_argument("empty graph");
    if (start[0] != 0) throw std::invalid_argument("face_start[0] must be 0");
    for (int i = 0; i < n; i++)
        if (start[i + 1] < start[i]) throw std::invalid_argument("face_start not ascending");
    for (int k = 0; k < start[n]; k++)
        if (nb[k] < -1 || nb[k] >= n) throw std::invalid_argument("neighbour out of range");
}

}  // namespace

std::vector<int> reverseSlots(int n, const std::vector<int>& start, const std::vector<int>& nb) {
    std::vector<int> rev(nb.size(), -1);
    for (int j = 0; j < n; j++)
        for (int k = start[j]; k < start[j + 1]; k++) {
            const int i = nb[k];
            if (i < 0) continue;
            for (int q = start[i]; q < start[i + 1]; q++)
                if (nb[q] == j) {
                    rev[k] = q;
                    break;
                }
        }
    return rev;
}

void NetGraphDev::build(int n, const std::vector<int>& start, const std::vector<int>& nb) {
    arena.release();
    g = NetGraph{};
    g.n = n;
    g.nPad = (n + 15) / 16 * 16;
    g.refs = (int)nb.size();
    g.start = arena.upload(start);
    g.nb = arena.upload(nb);
    g.rev = arena.upload(reverseSlots(n, start, nb));
    g.w = arena.zeros<float>(g.refs);
    g.wT = arena.zeros<float>(g.refs);
}

void NetWork::ensure(int nPad, int nLayers) {
    if (nPad <= rows && nLayers == layers) return;
    arena.release();
    rows = std::max(nPad, 16);
    layers = nLayers;
    const size_t R = (size_t)rows;
    const int blocks = (int)((R + 255) / 256);
    const size_t chunks = (R + kNetChunk - 1) / kNetChunk;
    r = arena.zeros<float>(R);
    y = arena.zeros<float>(R);
    dy = arena.zeros<float>(R);
    H0 = arena.zeros<uint16_t>(R * kNetIn);
    A.assign(layers, nullptr);
    H.assign(layers, nullptr);
    for (int l = 0; l < layers; l++) {
        A[l] = arena.zeros<uint16_t>(R * (l == 0 ? 2 * kNetIn : 2 * kNetHidden));
        H[l] = arena.zeros<uint16_t>(R * kNetHidden);
    }
    dH = arena.zeros<float>(R * kNetHidden);
    dA = arena.zeros<float>(R * 2 * kNetHidden);
    dZ = arena.zeros<uint16_t>(R * kNetHidden);
    partial = arena.zeros<float>(std::max(chunks * 2 * kNetHidden * kNetHidden, (size_t)blocks * kNetHidden));
    dpartial = arena.zeros<double>(2 * (size_t)blocks);
    x = arena.zeros<double>(R);
    Ax = arena.zeros<double>(R);
    scal = arena.zeros<double>(3);
    sigma = arena.zeros<double>(1);
    loss = arena.zeros<double>(1);
}

Mpnn::Mpnn(int layers, uint64_t seed, int device) : layers_(layers), device_(device) {
    if (layers < 1 || layers > 16) throw std::invalid_argument("layers must be in 1..16");
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    numParams_ = (int)offOut() + kNetHidden;
    // He-normal layers, readout N(0, 1/64)
    std::vector<float> p(numParams_);
    uint64_t s = seed;
    for (int l = 0; l < layers_; l++) {
        const double sd = std::sqrt(2.0 / K(l));
        for (size_t i = offW(l); i < offW(l + 1); i++) p[i] = (float)(sd * normal(s));
    }
    for (int f = 0; f < kNetHidden; f++) p[offOut() + f] = (float)(normal(s) / std::sqrt((double)kNetHidden));
    P_ = arena_.upload(p);
    G_ = arena_.zeros<float>(numParams_);
    M_ = arena_.zeros<float>(numParams_);
    V_ = arena_.zeros<float>(numParams_);
    for (int l = 0; l < layers_; l++) {
        Wb_.push_back(arena_.zeros<uint16_t>((size_t)K(l) * kNetHidden));
        WTb_.push_back(arena_.zeros<uint16_t>((size_t)K(l) * kNetHidden));
    }
    castWeights(stream_);
    AXS_HIP(hipStreamSynchronize(stream_));
}

Mpnn::~Mpnn() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    clearSamples();
    if (stream_) (void)hipStreamDestroy(stream_);
}

size_t Mpnn::offW(int l) const {
    size_t o = 0;
    for (int q = 0; q < l; q++) o += (size_t)K(q) * kNetHidden;
    return o;
}

void Mpnn::castWeights(hipStream_t s) {
    for (int l = 0; l < layers_; l++) netLaunchCastWeights(s, P_ + offW(l), K(l), kNetHidden, Wb_[l], WTb_[l]);
}

void Mpnn::getParams(float* out) const {
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamSynchronize(stream_));
    AXS_HIP(hipMemcpy(out, P_, (size_t)numParams_ * sizeof(float), hipMemcpyDeviceToHost));
}

void Mpnn::setParams(const float* in) {
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamSynchronize(stream_));
    AXS_HIP(hipMemcpy(P_, in, (size_t)numParams_ * sizeof(float), hipMemcpyHostToDevice));
    castWeights(stream_);
    AXS_HIP(hipStreamSynchronize(stream_));
}

int Mpnn::addSample(int n, const int* start, const int* nb, const double* AC, const double* AF, const double* b,
                    const double* target) {
    checkCSR(n, start, nb);
    if (!AC || !AF || !b || !target) throw std::invalid_argument("null sample array");
    AXS_HIP(hipSetDevice(device_));
    auto* smp = new Sample();
    try {
        const int S = start[n];
        std::vector<int> st(start, start + n + 1), nbv(nb, nb + S);
        smp->graph.build(n, st, nbv);
        double e = 0.0;
        for (int i = 0; i < n; i++) e += b[i] * target[i];
        if (!(e > 0.0)) throw std::invalid_argument("sample solution has no positive energy b.x");
        smp->estar = e;
        smp->AC = smp->arena.upload(std::vector<double>(AC, AC + n));
        smp->AF = smp->arena.upload(std::vector<double>(AF, AF + S));
        smp->b = smp->arena.upload(std::vector<double>(b, b + n));
    } catch (...) {
        delete smp;
        throw;
    }
    samples_.push_back(smp);
    return (int)samples_.size() - 1;
}

void Mpnn::clearSamples() {
    for (Sample* s : samples_) delete s;
    samples_.clear();
}

void Mpnn::forward(hipStream_t s, const NetGraph& g, const double* AC, const double* b, NetWork& w, bool scale) {
    if (scale) netLaunchScale(s, g.n, AC, b, w.dpartial, w.sigma);
    netLaunchFeatures(s, g, AC, b, w.sigma, w.r, w.H0);
    const uint16_t* in = w.H0;
    int D = kNetIn;
    for (int l = 0; l < layers_; l++) {
        netLaunchAggregate(s, g, in, D, w.A[l]);
        netLaunchGemm(s, w.A[l], Wb_[l], g.nPad, kNetHidden, 2 * D, nullptr, w.H[l]);
        in = w.H[l];
        D = kNetHidden;
    }
    netLaunchReadout(s, g.n, w.H[layers_ - 1], P_ + offOut(), w.y);
}

void Mpnn::backward(hipStream_t s, const NetGraph& g, const Sample& smp, NetWork& w) {
    netLaunchOutput(s, g.n, w.y, smp.AC, w.sigma, w.x);
    netLaunchEnergy(s, g, smp.AC, smp.AF, smp.b, w.x, w.Ax, w.dpartial, w.scal);
    netLaunchEnergyGrad(s, g.n, smp.b, w.Ax, smp.AC, w.sigma, w.scal, smp.estar, w.dy, w.loss);
    netLaunchReadoutBack(s, g.n, g.nPad, w.dy, w.H[layers_ - 1], P_ + offOut(), w.dH, w.partial, G_ + offOut());
    for (int l = layers_ - 1; l >= 0; l--) {
        netLaunchReluBack(s, (long)g.nPad * kNetHidden, w.dH, w.H[l], w.dZ);
        netLaunchGemmAtB(s, w.A[l], w.dZ, g.nPad, K(l), kNetHidden, w.partial, G_ + offW(l));
        if (l > 0) {
            netLaunchGemm(s, w.dZ, WTb_[l], g.nPad, K(l), kNetHidden, w.dA, nullptr);
            netLaunchAggregateBack(s, g, w.dA, kNetHidden, w.dH);
        }
    }
}

double Mpnn::lossGrad(int sample, float* grad) {
    if (sample < 0 || sample >= (int)samples_.size()) throw std::out_of_range("no such sample");
    AXS_HIP(hipSetDevice(device_));
    Sample& smp = *samples_[sample];
    const NetGraph& g = smp.graph.g;
    work_.ensure(g.nPad, layers_);
    netLaunchZero(stream_, numParams_, G_);
    netLaunchWeights(stream_, g, smp.AC, smp.AF);
    forward(stream_, g, smp.AC, smp.b, work_, true);
    backward(stream_, g, smp, work_);
    double loss = 0.0;
    AXS_HIP(hipMemcpyAsync(&loss, work_.loss, sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (grad) AXS_HIP(hipMemcpyAsync(grad, G_, (size_t)numParams_ * sizeof(float), hipMemcpyDeviceToHost, stream_));
    AXS_HIP(hipStreamSynchronize(stream_));
    return loss;
}

void Mpnn::train(int epochs, int batch, double lr, uint64_t seed, double* lossPerEpoch) {
    if (samples_.empty()) throw std::logic_error("no training samples");
    if (epochs <= 0 || batch <= 0 || !(lr > 0.0)) throw std::invalid_argument("bad training settings");
    AXS_HIP(hipSetDevice(device_));
    int maxPad = 0;
    for (Sample* s : samples_) maxPad = std::max(maxPad, s->graph.g.nPad);
    work_.ensure(maxPad, layers_);
    const int S = (int)samples_.size();
    Arena tmp;
    double* losses = tmp.zeros<double>(S);
    std::vector<int> order(S);
    std::iota(order.begin(), order.end(), 0);
    uint64_t rs = seed;
    std::vector<double> hl(S);
    for (int e = 0; e < epochs; e++) {
        for (int i = S - 1; i > 0; i--) std::swap(order[i], order[splitmix64(rs) % (uint64_t)(i + 1)]);
        for (int b0 = 0; b0 < S; b0 += batch) {
            netLaunchZero(stream_, numParams_, G_);
            for (int q = b0; q < std::min(S, b0 + batch); q++) {
                Sample& smp = *samples_[order[q]];
                const NetGraph& g = smp.graph.g;
                netLaunchWeights(stream_, g, smp.AC, smp.AF);
                forward(stream_, g, smp.AC, smp.b, work_, true);
                backward(stream_, g, smp, work_);
                AXS_HIP(hipMemcpyAsync(losses + order[q], work_.loss, sizeof(double), hipMemcpyDeviceToDevice,
                                       stream_));
            }
            step_++;
            netLaunchAdam(stream_, numParams_, P_, G_, M_, V_, (float)lr, 0.9f, 0.999f, 1e-8f, step_);
            castWeights(stream_);
        }
        AXS_HIP(hipMemcpyAsync(hl.data(), losses, S * sizeof(double), hipMemcpyDeviceToHost, stream_));
        AXS_HIP(hipStreamSynchronize(stream_));
        if (lossPerEpoch) lossPerEpoch[e] = std::accumulate(hl.begin(), hl.end(), 0.0) / S;
    }
    AXS_HIP(hipGetLastError());
}

void Mpnn::predict(int n, const int* start, const int* nb, const double* AC, const double* AF, const double* b,
                   double* x) {
    checkCSR(n, start, nb);
    if (!AC || !AF || !b || !x) throw std::invalid_argument("null array");
    AXS_HIP(hipSetDevice(device_));
    const int S = start[n];
    NetGraphDev gd;
    gd.build(n, std::vector<int>(start, start + n + 1), std::vector<int>(nb, nb + S));
    Arena tmp;
    const double* dAC = tmp.upload(std::vector<double>(AC, AC + n));
    const double* dAF = tmp.upload(std::vector<double>(AF, AF + S));
    const double* db = tmp.upload(std::vector<double>(b, b + n));
    double* dx = tmp.zeros<double>(n);
    work_.ensure(gd.g.nPad, layers_);
    enqueueForward(stream_, gd.g, dAC, dAF, db, dx, work_);
    AXS_HIP(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    AXS_HIP(hipStreamSynchronize(stream_));
    AXS_HIP(hipGetLastError());
}

void Mpnn::enqueueForward(hipStream_t s, const NetGraph& g, const double* AC, const double* AF, const double* b,
                          double* x, NetWork& w) {
    netLaunchWeights(s, g, AC, AF);
    forward(s, g, AC, b, w, true);
    netLaunchOutput(s, g.n, w.y, AC, w.sigma, x);
    // the energy-optimal multiple of the direction: never a worse start than 0
    netLaunchEnergy(s, g, AC, AF, b, x, w.Ax, w.dpartial, w.scal);
    netLaunchScaleVec(s, g.n, x, w.scal);
}

}  // namespace axs
