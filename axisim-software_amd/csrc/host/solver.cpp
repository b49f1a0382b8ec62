// solver.cpp -- host driver of the MI355X SIMPLE solver.
//
// Solver::runSimple (solver/source/solver.cpp:753-1396) restated for one GPU:
// the whole outer iteration -- assembly, 2 x linearMaxIter momentum sweeps,
// Rhie-Chow, the pressure-correction multigrid solve(s), the corrections and
// the residual norms -- is captured once as a hipGraph and replayed per
// iteration.  The convergence test runs on the device (kFinish), which raises
// ctl->done; the host launches graphs in growing batches and polls ctl between
// batches, so there is no per-iteration host synchronisation and no iteration
// runs past the one the reference would stop at.
#include "solver.hpp"

#include "net.hpp"

#include <algorithm>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>

namespace axs {

namespace {
constexpr int kSmoothBlockMax = 2048;     // cells a single workgroup smooths in LDS
constexpr int kTailMax = 1024;            // levels at most this small may run in the tail kernel
}

void Arena::release() {
    for (void* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
}

// ---------------------------------------------------------------------------
// MGDevice
// ---------------------------------------------------------------------------
void MGDevice::build(int n, const std::vector<int>& start, const std::vector<int>& nb, int maxLevels) {
    host = buildHierarchy(n, start, nb, maxLevels);
    arena.release();
    arena32.release();
    colorStart_.clear();
    dOrder_.clear();
    dColorStart_.clear();
    mixedReady = false;
    aug = inv = nullptr;
    invStatus = nullptr;
    if (ldsLimit_ == 0) ldsLimit_ = tailLdsLimit();
    tail64_ = selectTail(sizeof(double));
    tail32_ = TailPlan{};
    if (tail64_.lt >= 0) tail64_.image = arena.zeros<unsigned char>(tail64_.bytes);
    dev.assign(host.size(), DLevel{});
    aggThreadsHost_.assign(host.size(), {});
    for (size_t l = 0; l < host.size(); l++) {
        const Level& h = host[l];
        DLevel& d = dev[l];
        d.n = h.n;
        d.refs = h.refs();
        int minRow = 1 << 30;
        for (int i = 0; i < h.n; i++) {
            d.maxRow = std::max(d.maxRow, h.start[i + 1] - h.start[i]);
            minRow = std::min(minRow, h.start[i + 1] - h.start[i]);
        }
        d.uniformRow = (h.n > 0 && minRow == d.maxRow && h.start[0] == 0) ? d.maxRow : 0;
        if (!h.coarsest())
            for (size_t i = 0; i + 1 < h.memberStart.size(); i++)
                d.maxMembers = std::max(d.maxMembers, h.memberStart[i + 1] - h.memberStart[i]);
        if (tail64_.holds((int)l)) {
            // operator, iterates, CSR and maps inside the tail's image
            const TailLevel& e = tail64_.lv[l - tail64_.lt];
            double* v = (double*)tail64_.image;
            d.AC = v + e.AC;
            d.b = v + e.b;
            d.x = v + e.x0;
            d.xNew = v + e.x1;
            d.AF = v + e.AF;
            placeInts(tail64_, (int)l, d);
        } else {
            d.start = arena.upload(h.start);
            d.nb = arena.upload(h.nb);
            if (!h.coarsest()) {
                d.toCoarse = arena.upload(h.toCoarse);
                d.memberStart = arena.upload(h.memberStart);
                d.members = arena.upload(h.members);
            }
            if (l > 0) {
                d.AC = arena.zeros<double>(h.n);
                d.b = arena.zeros<double>(h.n);
                d.AF = arena.zeros<double>(h.refs());
                d.x = arena.zeros<double>(h.n);
                d.xNew = arena.zeros<double>(h.n);
            }
        }
        if (!h.coarsest()) {
            // workgroups of the fused residual + restriction: runs of whole
            // aggregates with at most kBlock members between them
            std::vector<int> agg{0};
            bool fits = true;
            for (int j = 0; j < host[l + 1].n; j++) {
                const int m = h.memberStart[j + 1] - h.memberStart[j];
                if (m > kBlock) fits = false;
                if (h.memberStart[j + 1] - h.memberStart[agg.back()] > kBlock) agg.push_back(j);
            }
            agg.push_back(host[l + 1].n);
            if (fits && host[l + 1].n > 0) {
                const Level& hc = host[l + 1];
                const int nB = (int)agg.size() - 1;
                std::vector<AggBlock> blocks(nB);
                std::vector<AggThread> threads((size_t)nB * kBlock, AggThread{-1, 0, 0, -1});
                for (int b = 0; b < nB; b++) {
                    const int j0 = agg[b], j1 = agg[b + 1];
                    const int p0 = h.memberStart[j0], np = h.memberStart[j1] - p0;
                    blocks[b] = AggBlock{j0, j1, hc.start[j0], hc.start[j1]};
                    AggThread* th = threads.data() + (size_t)b * kBlock;
                    for (int t = 0; t < np; t++) {
                        const int m = h.members[p0 + t];
                        th[t].m = m;
                        th[t].k0 = h.start[m];
                        th[t].ns = h.start[m + 1] - h.start[m];
                    }
                    for (int t = 0; t < j1 - j0; t++)
                        th[t].sum = (h.memberStart[j0 + t] - p0) | ((h.memberStart[j0 + t + 1] - p0) << 16);
                }
                d.aggBlocks = arena.upload(blocks);
                d.aggThreads = arena.upload(threads);
                aggThreadsHost_[l] = std::move(threads);
                d.nAggBlocks = nB;
            }
            d.nCoarse = host[l + 1].n;
            d.slotToCoarseSlot = arena.upload(h.slotToCoarseSlot);
            d.cslotStart = arena.upload(h.cslotStart);
            d.cslotFine = arena.upload(h.cslotFine);
            d.res = arena.zeros<double>(h.n);
        }
    }
    if (tail64_.lt >= 0) {
        tail64_.dev = arena.alloc<TailLevel>(tail64_.lv.size());
        AXS_HIP(hipMemcpy(tail64_.dev, tail64_.lv.data(), tail64_.lv.size() * sizeof(TailLevel),
                          hipMemcpyHostToDevice));
    }
    // fused level-0 sweeps: blocks of 1024 (else 512) cells with two halo
    // rings, when level 0 is large, its rows fit the kernel's registers and
    // every halo its LDS
    fuseL0_ = false;
    dl0_ = DSweepBlocks{};
    for (int cells : {1024, 512}) {
        if (fuseL0_ || host.size() < 2 || host[0].n < kFuseL0MinCells) break;
        buildSweepBlocks(host[0].n, host[0].start, host[0].nb, cells, 2, kSweepRegSlots, l0Blocks_);
        if (l0Blocks_.maxSlots <= kSweepRegSlots && l0Blocks_.maxHalo > sweepHaloMax(cells))
            buildSweepBlocks(host[0].n, host[0].start, host[0].nb, cells, 2, kSweepRegSlots, l0Blocks_, true);
        if (l0Blocks_.maxSlots <= kSweepRegSlots && l0Blocks_.maxHalo <= sweepHaloMax(cells)) {
            dl0_.cells = l0Blocks_.cells;
            dl0_.rings = l0Blocks_.rings;
            dl0_.nBlocks = l0Blocks_.nBlocks;
            dl0_.haloStart = arena.upload(l0Blocks_.haloStart);
            dl0_.ringEnd = arena.upload(l0Blocks_.ringEnd);
            dl0_.halo = arena.upload(l0Blocks_.halo.empty() ? std::vector<int>{0} : l0Blocks_.halo);
            dl0_.own = l0Blocks_.patches ? arena.upload(l0Blocks_.own) : nullptr;
            dl0_.locOwn = arena.upload(l0Blocks_.locOwn);
            dl0_.locHalo = arena.upload(l0Blocks_.locHalo.empty() ? std::vector<int16_t>{-1} : l0Blocks_.locHalo);
            fuseL0_ = true;
        }
    }
    // the level-1 fused restriction: every row's neighbours within its
    // workgroup's cells or a halo of at most kBlock cells, rows of at most
    // kL1Slots slots, members of at most 8
    fuseL1_ = false;
    l1f_ = L1Fused{};
    if (fuseL0_ && host.size() >= 3 && !host[1].coarsest() && dev[1].nAggBlocks > 0 && !tail64_.holds(1) &&
        dev[1].maxRow <= kL1Slots && dev[0].maxMembers <= 8) {
        const Level& h1 = host[1];
        const int nB = dev[1].nAggBlocks;
        const std::vector<AggThread>& th = aggThreadsHost_[1];
        std::vector<int> halo((size_t)nB * kBlock, -1);
        std::vector<int16_t> loc((size_t)nB * kBlock * kL1Slots, -1);
        std::vector<int> local(h1.n, -1), touched;
        bool ok = true;
        for (int b = 0; b < nB && ok; b++) {
            touched.clear();
            for (int t = 0; t < kBlock; t++) {
                const int m = th[(size_t)b * kBlock + t].m;
                if (m >= 0) {
                    local[m] = t;
                    touched.push_back(m);
                }
            }
            int nh = 0;
            for (int t = 0; t < kBlock && ok; t++) {
                const int m = th[(size_t)b * kBlock + t].m;
                if (m < 0) continue;
                for (int k = h1.start[m], q = 0; k < h1.start[m + 1]; k++, q++) {
                    const int o = h1.nb[k];
                    if (o < 0) continue;
                    if (local[o] < 0) {
                        if (nh == kBlock) {
                            ok = false;
                            break;
                        }
                        local[o] = kBlock + nh;
                        halo[(size_t)b * kBlock + nh++] = o;
                        touched.push_back(o);
                    }
                    loc[((size_t)b * kBlock + t) * kL1Slots + q] = (int16_t)local[o];
                }
            }
            for (int c : touched) local[c] = -1;
        }
        if (ok) {
            l1f_.halo = arena.upload(halo);
            l1f_.loc = arena.upload(loc);
            fuseL1_ = true;
        }
    }
}

// Upload level l's CSR and transfer maps into the int part of the tail image
// and point the level at them.
void MGDevice::placeInts(const TailPlan& t, int l, DLevel& d) {
    const Level& h = host[l];
    const TailLevel& e = t.lv[l - t.lt];
    int* ip = (int*)(t.image + t.intBase);
    auto put = [&](int off, const std::vector<int>& src) {
        if (!src.empty())
            AXS_HIP(hipMemcpy(ip + off, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
        return ip + off;
    };
    d.start = put(e.start, h.start);
    d.nb = put(e.nb, h.nb);
    if (e.perm >= 0) {
        // the coarsest level's sweep order of threads: cells by slot count
        // (stable), so each wave's rows are about equally long
        std::vector<int> perm(h.n);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int a, int b) { return h.start[a + 1] - h.start[a] < h.start[b + 1] - h.start[b]; });
        put(e.perm, perm);
    }
    if (!h.coarsest()) {
        d.toCoarse = put(e.toC, h.toCoarse);
        d.memberStart = put(e.mS, h.memberStart);
        d.members = put(e.mem, h.members);
    }
}

// LDS image of levels lt..L-1 for the tail kernel: values (AC, b, x ping-pong,
// AF) first, then ints (CSR, and the transfer maps of every level but the last).
// The tail levels' device arrays live in a global copy of exactly this image,
// so the kernel stages it with one contiguous copy.
MGDevice::TailPlan MGDevice::planTail(int lt, size_t vsize) const {
    TailPlan t;
    const int L = (int)host.size();
    t.lt = lt;
    t.nl = L - lt;
    // Values the kernel reads before writing them (every level's AC and AF,
    // level lt's b and x0) come first, then the integer CSR and maps, then
    // the values it only writes before reading (x1, the deeper levels' b and
    // x0): a launch stages the image up to the end of the integers only.
    long tv = 0, ti = 0;
    int maxN = 0;
    for (int l = lt; l < L; l++) {
        const Level& h = host[l];
        const bool last = l == L - 1;
        TailLevel e{};
        e.n = h.n;
        e.refs = h.refs();
        e.nc = last ? 0 : host[l + 1].n;
        e.AC = (int)tv; tv += h.n;
        if (l == lt) {
            e.b = (int)tv; tv += h.n;
            e.x0 = (int)tv; tv += h.n;
        }
        e.AF = (int)tv; tv += e.refs;
        e.start = (int)ti; ti += h.n + 1;
        e.nb = (int)ti; ti += e.refs;
        if (!last) {
            e.toC = (int)ti; ti += h.n;
            e.mS = (int)ti; ti += e.nc + 1;
            e.mem = (int)ti; ti += h.n;
        }
        e.perm = -1;
        if (last && h.n > 64) {
            e.perm = (int)ti;
            ti += h.n;
        }
        t.lv.push_back(e);
        maxN = std::max(maxN, h.n);
    }
    t.intBase = (int)(((size_t)tv * vsize + 15) / 16 * 16);
    t.staged = ((size_t)t.intBase + (size_t)ti * sizeof(int) + 15) / 16 * 16;
    // the write-first values after the integers, in value units from the start
    long tw = (long)(t.staged / vsize);
    for (int l = lt; l < L; l++) {
        TailLevel& e = t.lv[l - lt];
        if (l != lt) {
            e.b = (int)tw; tw += e.n;
            e.x0 = (int)tw; tw += e.n;
        }
        e.x1 = (int)tw; tw += e.n;
    }
    t.bytes = ((size_t)tw * vsize + 15) / 16 * 16;
    t.threads = std::min(1024, std::max(64, (maxN + 63) / 64 * 64));
    return t;
}

// The tail starts at the first level below level 0 (whose arrays belong to the
// caller) that is small enough and whose LDS image, with every coarser level,
// fits one workgroup.
MGDevice::TailPlan MGDevice::selectTail(size_t vsize) const {
    const int L = (int)host.size();
    const size_t budget = ldsLimit_ > kTailStaticLds ? (size_t)(ldsLimit_ - kTailStaticLds) : 0;
    for (int l = 1; l < L; l++) {
        if (host[l].n > kTailMax || L - l > kTailLevels) continue;
        TailPlan c = planTail(l, vsize);
        if (c.bytes <= budget) return c;
    }
    return TailPlan{};
}

TailLaunch MGDevice::TailPlan::launch() const {
    TailLaunch l;
    l.plan = dev;
    l.image = image;
    l.vecs = (int)(staged / 16);
    l.lt = lt;
    l.nl = nl;
    l.intBase = intBase;
    l.threads = threads;
    l.ldsBytes = bytes;
    return l;
}

// bytes the tail kernel moves between HBM and LDS: the image in, level lt's x out
double MGDevice::tailTrafficBytes(const TailPlan& t, size_t vsize) const {
    if (t.lt < 0) return 0.0;
    return (double)t.staged + (double)vsize * t.lv[0].n;
}

void MGDevice::bindLevel0(double* AC, double* b, double* AF, double* x, double* xNew) {
    dev[0].AC = AC;
    dev[0].b = b;
    dev[0].AF = AF;
    dev[0].x = x;
    dev[0].xNew = xNew;
    x0Home_ = x;
    x0Other_ = xNew;
}

void MGDevice::swapX(int l) { std::swap(dev[l].x, dev[l].xNew); }

namespace {
size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

size_t MGDevice::workBytes() const {
    size_t bytes = 0;
    const int L = (int)host.size();
    for (int l = 0; l < L; l++) {
        const Level& h = host[l];
        const size_t n = h.n, r = h.refs();
        if (!h.coarsest()) bytes += align256(n * sizeof(double));  // res (tail levels' too)
        if (tail64_.holds(l)) continue;
        if (l > 0) bytes += 4 * align256(n * sizeof(double)) + align256(r * sizeof(double));  // AC b x xNew, AF
        else bytes += align256(n * sizeof(double));                                          // xNew
    }
    if (tail64_.lt >= 0) bytes += align256(tail64_.bytes);
    return bytes;
}

MGDevice::Bound MGDevice::bindWork(unsigned char* base, const double* AC, const double* b, const double* AF, double* x,
                                   hipStream_t s) {
    Bound saved{dev, tail64_.image, x0Home_, x0Other_};
    size_t off = 0;
    auto take = [&](size_t count) {
        double* p = (double*)(base + off);
        off += align256(count * sizeof(double));
        return p;
    };
    unsigned char* tailBase = nullptr;
    const int L = (int)host.size();
    for (int l = 0; l < L; l++) {
        const Level& h = host[l];
        DLevel& d = dev[l];
        if (!h.coarsest()) d.res = take(h.n);
        if (tail64_.holds(l)) continue;
        if (l > 0) {
            d.AC = take(h.n);
            d.b = take(h.n);
            d.x = take(h.n);
            d.xNew = take(h.n);
            d.AF = take(h.refs());
        } else {
            d.AC = const_cast<double*>(AC);
            d.b = const_cast<double*>(b);
            d.AF = const_cast<double*>(AF);
            d.x = x;
            d.xNew = take(h.n);
        }
    }
    if (tail64_.lt >= 0) {
        tailBase = base + off;
        off += align256(tail64_.bytes);
        // the tail levels' arrays are offsets into the image: move every pointer
        // that lies inside it by the same amount
        unsigned char* img = tail64_.image;
        const size_t n = tail64_.bytes;
        auto move = [&](auto*& p) {
            using T = std::remove_reference_t<decltype(*p)>;
            unsigned char* q = (unsigned char*)p;
            if (q >= img && q < img + n) p = (T*)(tailBase + (q - img));
        };
        for (int l = tail64_.lt; l < L; l++) {
            DLevel& d = dev[l];
            move(d.AC); move(d.b); move(d.AF); move(d.x); move(d.xNew);
            move(d.start); move(d.nb); move(d.toCoarse); move(d.memberStart); move(d.members);
        }
        // the integer maps the kernel stages (the values are written before read)
        AXS_HIP(hipMemcpyAsync(tailBase + tail64_.intBase, img + tail64_.intBase, tail64_.staged - tail64_.intBase,
                               hipMemcpyDeviceToDevice, s));
        tail64_.image = tailBase;
    }
    x0Home_ = x;
    x0Other_ = dev[0].xNew;
    return saved;
}

void MGDevice::unbindWork(const Bound& saved) {
    dev = saved.dev;
    tail64_.image = saved.tailImage;
    x0Home_ = saved.x0Home;
    x0Other_ = saved.x0Other;
}

// smoothen / smoothenRegular / smoothenSingleBlock, multigrid.cu:664-725
void MGDevice::smooth(hipStream_t s, int l, int sweeps, const MGParams& p, const Ctl* ctl) {
    DLevel& L = dev[l];
    if (L.n <= 0 || sweeps <= 0) return;
    if (sweeps > 1 && L.n <= kSmoothBlockMax) {
        launchMGSmoothBlock(s, L, p.weight, sweeps, ctl);
        return;
    }
    for (int i = 0; i < sweeps; i++) {
        launchMGJacobi(s, L, p.weight, ctl);
        swapX(l);
    }
}

// computeResidual + buildRestrictionKernel + the coarse x memset
// (multigrid.cu:641-655), with the coarse level's first sweep from that zero
// folded in when sweep1 is set.
void MGDevice::restrictTo(hipStream_t s, int l, int sweep1, bool fp32, const MGParams& p, const Ctl* ctl) {
    if (dev[l].nAggBlocks > 0) {
        launchMGResidualRestrict(s, dev[l], dev[l + 1], p.weight, sweep1, fp32, ctl);
        return;
    }
    if (fp32) launchMGResidual32(s, dev[l], ctl);
    else launchMGResidual(s, dev[l], ctl);
    launchMGRestrictSum(s, dev[l], dev[l + 1], p.weight, sweep1, fp32, ctl);
}

// vCycle, multigrid.cu:607-662.  preDone: pre-smoothing sweeps of level l
// already applied (by the restriction above it).
void MGDevice::vcycle(hipStream_t s, int l, int preDone, const MGParams& p, const Ctl* ctl) {
    // coarse_direct: as in vcycle32, the inverse applied inside the tail when
    // one workgroup formed it, else across the GPU after ordinary launches
    const bool wideDirect = p.coarseDirect && coarsestCells() > kDirectOneBlock;
    if (l == tail64_.lt && !wideDirect) {
        launchMGTail(s, tail64_.launch(), p.weight, preDone, p.pre, p.post, p.coarseSweeps, false,
                     p.coarseDirect ? inv : nullptr, ctl);
        return;
    }
    if (host[l].coarsest()) {
        if (p.coarseDirect) launchCoarseApply64(s, dev[l], inv, ctl);
        else smooth(s, l, p.coarseSweeps - preDone, p, ctl);
        return;
    }
    smooth(s, l, p.pre - preDone, p, ctl);
    const int s1 = host[l + 1].coarsest() ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1);
    restrictTo(s, l, s1, false, p, ctl);
    vcycle(s, l + 1, s1, p, ctl);
    if (p.post >= 1) {
        launchMGProlongSmooth(s, dev[l], dev[l + 1], p.weight, ctl);
        swapX(l);
        smooth(s, l, p.post - 1, p, ctl);
    } else {
        launchMGProlong(s, dev[l], dev[l + 1], ctl);
    }
}

void MGDevice::swapX32(int l) { std::swap(dev[l].x32, dev[l].xNew32); }

bool MGDevice::fusesLevel0(const MGParams& p) const {
    return fuseL0_ && !fuseL0Off_ && host.size() >= 2 && p.pre == 1 && p.post == 1 &&
           p.smoother == AXS_MG_SMOOTH_JACOBI;
}

// The V-cycles of one solve with the level-0 work fused across cycles: each
// cycle's pre-sweep and residual run in one launch together with the previous
// cycle's prolongation and post-sweep; the last cycle's post-sweep runs alone.
// The same sweeps and residuals as vcycle(0) repeated, bit for bit.  fp32: the
// same on the FP32 copies, as vcycle32(0) repeated, the last post-sweep also
// adding the iterate into `apply` (FP64) when given.
void MGDevice::cyclesFused(hipStream_t s, int cycles, const MGParams& p, const Ctl* ctl, bool fp32, double* apply) {
    // a coarsest level solved directly takes no sweep from zero
    const bool direct = p.coarseDirect != 0;
    auto sweep1Of = [&](int l) { return host[l].coarsest() ? (!direct && p.coarseSweeps >= 1) : (p.pre >= 1); };
    const int s1 = sweep1Of(1);
    // level 1 folded into the restriction when no pre-sweep runs there between
    const bool l1 = fuseL1_ && s1 == 1 && p.pre == 1 && !(fp32 && tail32_.holds(1));
    const int s2 = host.size() < 3 ? 0 : sweep1Of(2);
    for (int c = 0; c < cycles; c++) {
        // FP32: the first launch starts from e = 0 on b32 (x = 0) or on the
        // FP64 defect it forms itself
        if (fp32 && c == 0) launchMGLevel0BlockStart32(s, dev[0], dev[1], dl0_, p.weight, !p.startZero, ctl);
        else launchMGLevel0Block(s, dev[0], dev[1], dl0_, c == 0 ? 1 : 2, c > 0, p.weight, fp32, ctl);
        fp32 ? swapX32(0) : swapX(0);
        if (l1) {
            // vcycle(1) with its restriction taken by the fused launch
            launchMGRestrictFused01(s, dev[0], dev[1], dev[2], l1f_, p.weight, s1, s2, fp32, ctl);
            if (fp32) {
                vcycle32(s, 2, s2, p, ctl);
                launchMGProlongSmooth32(s, dev[1], dev[2], (float)p.weight, ctl);
                swapX32(1);
                smooth32(s, 1, p.post - 1, p, ctl);
            } else {
                vcycle(s, 2, s2, p, ctl);
                launchMGProlongSmooth(s, dev[1], dev[2], p.weight, ctl);
                swapX(1);
                smooth(s, 1, p.post - 1, p, ctl);
            }
            continue;
        }
        launchMGRestrictSum(s, dev[0], dev[1], p.weight, s1, fp32, ctl);
        if (fp32) vcycle32(s, 1, s1, p, ctl);
        else vcycle(s, 1, s1, p, ctl);
    }
    if (fp32) {
        launchMGProlongSmooth32(s, dev[0], dev[1], (float)p.weight, ctl, apply);
        swapX32(0);
    } else {
        launchMGProlongSmooth(s, dev[0], dev[1], p.weight, ctl);
        swapX(0);
    }
}

// Greedy colouring of every level's graph (buildMeshColoring order), built on
// first use of the Gauss-Seidel smoother.
void MGDevice::ensureColoring() {
    if (colorStart_.size() == host.size()) return;
    colorStart_.assign(host.size(), {});
    dOrder_.assign(host.size(), nullptr);
    dColorStart_.assign(host.size(), nullptr);
    for (size_t l = 0; l < host.size(); l++) {
        std::vector<int> order;
        greedyColoring(host[l].n, host[l].start.data(), host[l].nb.data(), colorStart_[l], order);
        dOrder_[l] = arena.upload(order);
        dColorStart_[l] = arena.upload(colorStart_[l]);
    }
}

void MGDevice::smoothGS(hipStream_t s, int l, int sweeps, bool fp32, const Ctl* ctl) {
    const DLevel& L = dev[l];
    if (L.n <= 0 || sweeps <= 0) return;
    const int nColors = (int)colorStart_[l].size() - 1;
    if (L.n <= kSmoothBlockGSMax) {
        launchMGSmoothBlockGS(s, L, dOrder_[l], dColorStart_[l], nColors, sweeps, fp32, ctl);
        return;
    }
    for (int it = 0; it < sweeps; it++)
        for (int c = 0; c < nColors; c++) {
            const int begin = colorStart_[l][c];
            launchMGGSColor(s, L, dOrder_[l], begin, colorStart_[l][c + 1] - begin, fp32, ctl);
        }
}

// vCycle (multigrid.cu:607-662) with Gauss-Seidel smoothing: x stays in place
// on every level, the prolongation is a plain x += P x_c.
void MGDevice::vcycleGS(hipStream_t s, int l, bool fp32, const MGParams& p, const Ctl* ctl) {
    if (host[l].coarsest()) {
        if (p.coarseDirect && fp32) launchCoarseApply(s, dev[l], inv, ctl);
        else if (p.coarseDirect) launchCoarseApply64(s, dev[l], inv, ctl);
        else smoothGS(s, l, p.coarseSweeps, fp32, ctl);
        return;
    }
    smoothGS(s, l, p.pre, fp32, ctl);
    restrictTo(s, l, 0, fp32, p, ctl);
    vcycleGS(s, l + 1, fp32, p, ctl);
    if (fp32) launchMGProlong32(s, dev[l], dev[l + 1], ctl);
    else launchMGProlong(s, dev[l], dev[l + 1], ctl);
    smoothGS(s, l, p.post, fp32, ctl);
}

void MGDevice::ensureMixed() {
    if (mixedReady) return;
    arena32.release();
    tail32_ = selectTail(sizeof(float));
    if (tail32_.lt >= 0) tail32_.image = arena32.zeros<unsigned char>(tail32_.bytes);
    for (size_t l = 0; l < host.size(); l++) {
        DLevel& d = dev[l];
        if (tail32_.holds((int)l)) {
            // FP32 copies inside the FP32 tail image, with its own copy of the
            // CSR and maps (the FP64 arrays keep theirs)
            const TailLevel& e = tail32_.lv[l - tail32_.lt];
            float* v = (float*)tail32_.image;
            d.AC32 = v + e.AC;
            d.b32 = v + e.b;
            d.x32 = v + e.x0;
            d.xNew32 = v + e.x1;
            d.AF32 = v + e.AF;
            DLevel scratch = d;
            placeInts(tail32_, (int)l, scratch);
        } else {
            d.AC32 = arena32.zeros<float>(d.n);
            d.AF32 = arena32.zeros<float>(d.refs);
            d.b32 = arena32.zeros<float>(d.n);
            d.x32 = arena32.zeros<float>(d.n);
            d.xNew32 = arena32.zeros<float>(d.n);
        }
        if (!host[l].coarsest()) d.res32 = arena32.zeros<float>(d.n);
    }
    if (tail32_.lt >= 0) {
        tail32_.dev = arena32.alloc<TailLevel>(tail32_.lv.size());
        AXS_HIP(hipMemcpy(tail32_.dev, tail32_.lv.data(), tail32_.lv.size() * sizeof(TailLevel),
                          hipMemcpyHostToDevice));
    }
    const int nc = coarsestCells();
    if (nc <= kDirectMax) {
        const size_t ld = (size_t)coarseInverseLd(nc);
        aug = arena32.zeros<double>(coarseInverseWork(nc));
        inv = arena32.zeros<double>(ld * ld);
        invStatus = arena32.zeros<int>(1);
    }
    mixedReady = true;
}

void MGDevice::ensureInverse() {
    if (inv) return;
    const int nc = coarsestCells();
    if (nc > kDirectMax) return;  // the solve rejects it
    const size_t ld = (size_t)coarseInverseLd(nc);
    aug = arena32.zeros<double>(coarseInverseWork(nc));
    inv = arena32.zeros<double>(ld * ld);
    invStatus = arena32.zeros<int>(1);
}

void MGDevice::smooth32(hipStream_t s, int l, int sweeps, const MGParams& p, const Ctl* ctl) {
    DLevel& L = dev[l];
    if (L.n <= 0 || sweeps <= 0) return;
    if (sweeps > 1 && L.n <= kSmoothBlockMax) {
        launchMGSmoothBlock32(s, L, (float)p.weight, sweeps, ctl);
        return;
    }
    for (int i = 0; i < sweeps; i++) {
        launchMGJacobi32(s, L, (float)p.weight, ctl);
        swapX32(l);
    }
}

void MGDevice::vcycle32(hipStream_t s, int l, int preDone, const MGParams& p, const Ctl* ctl, double* apply) {
    // a coarsest level solved directly that is beyond one workgroup's inverse
    // is applied across the GPU (one wave per row), not inside the tail: the
    // tail's levels then run as ordinary launches
    const bool wideDirect = p.coarseDirect && coarsestCells() > kDirectOneBlock;
    if (l == tail32_.lt && !wideDirect) {
        launchMGTail(s, tail32_.launch(), p.weight, preDone, p.pre, p.post, p.coarseSweeps, true,
                     p.coarseDirect ? inv : nullptr, ctl);
        return;
    }
    if (host[l].coarsest()) {
        if (p.coarseDirect) launchCoarseApply(s, dev[l], inv, ctl);
        else smooth32(s, l, p.coarseSweeps - preDone, p, ctl);
        return;
    }
    smooth32(s, l, p.pre - preDone, p, ctl);
    const int s1 = host[l + 1].coarsest() ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1);
    restrictTo(s, l, s1, true, p, ctl);
    vcycle32(s, l + 1, s1, p, ctl);
    if (p.post >= 1) {
        launchMGProlongSmooth32(s, dev[l], dev[l + 1], (float)p.weight, ctl, apply);
        swapX32(l);
        smooth32(s, l, p.post - 1, p, ctl);
    } else {
        launchMGProlong32(s, dev[l], dev[l + 1], ctl);
    }
}

// Mixed precision: every cycle is x += B (b - A x), the FP64 defect feeding an
// FP32 V-cycle started from zero (the same affine map as the FP64 cycle).
void MGDevice::enqueueSolveMixed(hipStream_t s, const MGParams& p, const Ctl* ctl, bool buildOperators) {
    ensureMixed();
    const int L = (int)host.size();
    if (p.coarseDirect && coarsestCells() > kDirectMax)
        throw std::domain_error("coarsest level too large for the dense FP64 solve");
    if (buildOperators) enqueueOperators(s, p, ctl);
    const int cycles = p.cycles > 1 ? p.cycles : 1;
    // the first sweep from e = 0 is fused into the defect kernel
    const bool single = L == 1;
    if (p.smoother == AXS_MG_SMOOTH_RBGS) {
        ensureColoring();
        for (int c = 0; c < cycles; c++) {
            launchMGDefect(s, dev[0], p.weight, 0, ctl);
            vcycleGS(s, 0, true, p, ctl);
            launchMGApply(s, dev[0], ctl);
        }
        return;
    }
    const int first = (single ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1)) ? 1 : 0;
    // x += e rides on level 0's post-sweep when that sweep is the cycle's last
    const bool fuseApply = !single && tail32_.lt != 0 && p.post == 1;
    if (p.refine == AXS_REFINE_SOLVE) {
        // one FP64 defect, then every cycle in FP32 on the correction e (from
        // e = 0), x += e once at the end; fused, the defect is formed by the
        // first level-0 launch (or is b itself when x = 0)
        if (fusesLevel0(p) && (tail32_.lt < 0 || tail32_.lt >= 2)) {
            cyclesFused(s, cycles, p, ctl, true, dev[0].x);
            return;
        }
        launchMGDefect(s, dev[0], p.weight, 0, ctl);
        for (int c = 0; c < cycles; c++) vcycle32(s, 0, 0, p, ctl, (fuseApply && c + 1 == cycles) ? dev[0].x : nullptr);
        if (!fuseApply) launchMGApply(s, dev[0], ctl);
        return;
    }
    for (int c = 0; c < cycles; c++) {
        launchMGDefect(s, dev[0], p.weight, first, ctl);
        vcycle32(s, 0, first, p, ctl, fuseApply ? dev[0].x : nullptr);
        if (!fuseApply) launchMGApply(s, dev[0], ctl);
    }
}

void MGDevice::enqueueOperators(hipStream_t s, const MGParams& p, const Ctl* ctl) {
    if (p.precision == AXS_PREC_MIXED) {
        ensureMixed();
        if (!p.l0Copies) launchMGCast(s, dev[0], ctl);
        buildOperators(s, ctl);  // also writes the FP32 copies
        if (p.coarseDirect && p.formInverse) launchCoarseInvert(s, dev[host.size() - 1], aug, inv, invStatus, ctl);
    } else {
        buildOperators(s, ctl);
        if (p.coarseDirect && p.formInverse) {
            ensureInverse();
            launchCoarseInvert(s, dev[host.size() - 1], aug, inv, invStatus, ctl);
        }
    }
}

// Galerkin coarse operators of every level (buildCoarseOperator,
// multigrid.cu:350-389), each in two coalesced passes.
void MGDevice::buildOperators(hipStream_t s, const Ctl* ctl) {
    for (size_t l = 1; l < dev.size(); l++) {
        if (dev[l - 1].nAggBlocks > 0) {
            launchMGCoarseOperator(s, dev[l - 1], dev[l], ctl);
            continue;
        }
        launchMGCoarseDiag(s, dev[l - 1], ctl);
        launchMGCoarseGather(s, dev[l - 1], dev[l], ctl);
    }
}

void MGDevice::enqueueSolve(hipStream_t s, const MGParams& p, const Ctl* ctl, bool buildOperators) {
    if (p.precision == AXS_PREC_MIXED) {
        enqueueSolveMixed(s, p, ctl, buildOperators);
        return;
    }
    if (p.coarseDirect) {
        if (coarsestCells() > kDirectMax) throw std::domain_error("coarsest level too large for the dense FP64 solve");
        ensureInverse();
    }
    if (buildOperators) enqueueOperators(s, p, ctl);
    const int cycles = p.cycles > 1 ? p.cycles : 1;  // cycleCount, multigrid.cuh:226
    if (p.smoother == AXS_MG_SMOOTH_RBGS) {
        ensureColoring();
        for (int c = 0; c < cycles; c++) vcycleGS(s, 0, false, p, ctl);
        return;  // x never leaves the level-0 buffer bound as x
    }
    if (fusesLevel0(p)) cyclesFused(s, cycles, p, ctl);
    else
        for (int c = 0; c < cycles; c++) vcycle(s, 0, 0, p, ctl);
    if (dev[0].x != x0Home_) {
        launchCopy(s, x0Home_, dev[0].x, dev[0].n, ctl);
        dev[0].x = x0Home_;
        dev[0].xNew = x0Other_;
    }
}

// Algorithmic bytes of one solve (distinct elements read + written, FP64 / int32).
// Per launch: distinct elements read + written, values of size v (8 or 4),
// int32 indices 4 B.  Mirrors the launches enqueueSolve makes.
// Residual + restriction of level l: fine rows (AC, b, x, AF/nb) and their
// range (start + members, or the fused kernel's 16-byte thread entry), the
// residual written and re-read unless fused, memberStart unless fused, coarse
// b and x.
double MGDevice::restrictBytes(int l, double v) const {
    const double n = host[l].n, S = host[l].refs(), nc = host[l + 1].n;
    if (dev[l].nAggBlocks > 0) return (3.0 * v + 16.0) * n + (v + 4.0) * S + 2.0 * v * nc;
    return (5.0 * v + 8.0) * n + (v + 4.0) * S + (2.0 * v + 4.0) * nc;
}

double MGDevice::bytesPerSolve(const MGParams& p, bool buildOperators) const {
    const bool mixed = p.precision == AXS_PREC_MIXED;
    const double v = mixed ? 4.0 : 8.0;
    const int L = (int)host.size();
    const TailPlan& tail = mixed ? tail32_ : tail64_;
    double bytes = 0.0;
    if (buildOperators) {
        if (mixed) bytes += 12.0 * host[0].n + 12.0 * host[0].refs();  // FP32 cast of level 0
        const double w32 = mixed ? 4.0 : 0.0;                         // FP32 copies written
        for (int l = 1; l < L; l++) {
            const Level& f = host[l - 1];
            const Level& c = host[l];
            const double X = (double)f.cslotFine.size();
            // diagonals (AC and the row range -- start, or the fused kernel's
            // 16-byte thread entry -- nb/AF/slot map; one diagonal per fine cell
            // written and re-read unless fused) + ordered gathers + coarse writes
            bytes += (dev[l - 1].nAggBlocks > 0 ? 24.0 : 32.0) * f.n + 16.0 * f.refs() + 12.0 * X +
                     (12.0 + w32) * c.n + (12.0 + w32) * c.refs();
        }
        if (p.coarseDirect) bytes += 8.0 * 3.0 * host.back().n * host.back().n;
    }
    if (p.smoother == AXS_MG_SMOOTH_RBGS) {
        // a colour sweep pass reads AC, b, start, the colour order and the row (AF, nb)
        // and updates x in place; a level held in one workgroup is read once
        auto gs = [&](int l, int sweeps) {
            if (sweeps <= 0) return 0.0;
            const double pass = (4.0 * v + 8.0) * host[l].n + (v + 4.0) * host[l].refs();
            return host[l].n <= kSmoothBlockGSMax ? pass : sweeps * pass;
        };
        double cyc = mixed ? 32.0 * host[0].n + 12.0 * host[0].refs() + 8.0 * host[0].n + 20.0 * host[0].n : 0.0;
        for (int l = 0; l < L; l++) {
            const double n = host[l].n;
            if (host[l].coarsest()) {
                cyc += p.coarseDirect ? 8.0 * n * n + 2.0 * v * n : gs(l, p.coarseSweeps);
                break;
            }
            const double nc = host[l + 1].n;
            cyc += gs(l, p.pre) + gs(l, p.post);
            // residual + restriction (the residual written and re-read unless fused)
            cyc += restrictBytes(l, v);
            cyc += (2.0 * v + 4.0) * n + v * nc;                                  // prolongation
        }
        return bytes + (p.cycles > 1 ? p.cycles : 1) * cyc;
    }
    auto jac = [&](int l) { return (4.0 * v + 4.0) * host[l].n + (v + 4.0) * host[l].refs(); };
    auto sm = [&](int l, int sweeps) {
        if (sweeps <= 0) return 0.0;
        if (sweeps > 1 && host[l].n <= kSmoothBlockMax) return (4.0 * v + 4.0) * host[l].n + (v + 4.0) * host[l].refs();
        return sweeps * jac(l);
    };
    // one cycle from level `from` down, its pre-smoothing partly done
    auto cycleFrom = [&](int from, int preDone) {
        double cyc = 0.0;
        for (int l = from; l < L; l++) {
            if (l == tail.lt) {
                cyc += tailTrafficBytes(tail, (size_t)v);
                break;
            }
            if (host[l].coarsest()) {
                cyc += p.coarseDirect ? 8.0 * host[l].n * host[l].n + 2.0 * v * host[l].n
                                                 : sm(l, p.coarseSweeps - preDone);
                break;
            }
            const double n = host[l].n, S = host[l].refs(), nc = host[l + 1].n;
            cyc += sm(l, p.pre - preDone);
            const int s1 =
                host[l + 1].coarsest() ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1);
            // restriction (+ coarse AC for the folded sweep)
            cyc += restrictBytes(l, v) + (s1 ? v * nc : 0.0);
            if (p.post >= 1) cyc += (4.0 * v + 8.0) * n + (v + 4.0) * S + v * nc + sm(l, p.post - 1);
            else cyc += (2.0 * v + 4.0) * n + v * nc;
            preDone = s1;
        }
        return cyc;
    };
    const int cycles = p.cycles > 1 ? p.cycles : 1;
    const bool once = mixed && p.refine == AXS_REFINE_SOLVE;
    // the FP64 defect of a solve refined once (AC, b, x, start + AF/nb; b32 and
    // x32 written) and its x += e (x read and written)
    const double defect = 40.0 * host[0].n + 12.0 * host[0].refs(), applyX = 16.0 * host[0].n;
    if (fusesLevel0(p) && (!mixed || (once && (tail32_.lt < 0 || tail32_.lt >= 2)))) {
        // level 0 as cyclesFused launches it: (sweep + residual), then per later
        // cycle (prolongation + 2 sweeps + residual) -- rows AC, b, x, start,
        // AF/nb (+ toCoarse, coarse x), x and the residual written -- the
        // restriction per cycle (residual, members; coarse b, x, AC), and the
        // last post-sweep
        const double n = host[0].n, S = host[0].refs(), nc = host[1].n;
        const int s1 = host[1].coarsest() ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1);
        double l0 = (5.0 * v + 4.0) * n + (v + 4.0) * S + (once ? defect + applyX : 0.0);
        l0 += (cycles - 1) * ((5.0 * v + 8.0) * n + (v + 4.0) * S + v * nc);
        l0 += cycles * ((v + 4.0) * n + (2.0 * v + 4.0) * nc + (s1 ? v * nc : 0.0));
        l0 += (4.0 * v + 8.0) * n + (v + 4.0) * S + v * nc;
        // with level 1 folded into the restriction, its b and x come from LDS,
        // not from a second read
        const double l1 = (fuseL1_ && s1 == 1 && p.pre == 1 && !(mixed && tail32_.holds(1))) ? 2.0 * v * nc : 0.0;
        return bytes + l0 + cycles * (cycleFrom(1, s1) - l1);
    }
    if (once) {
        const bool fuseApply = L > 1 && tail32_.lt != 0 && p.post == 1;
        return bytes + defect + cycles * cycleFrom(0, 0) + (fuseApply ? applyX : 20.0 * host[0].n);
    }
    double cyc = 0.0;
    int preDone = 0;
    if (mixed) {
        // FP64 defect (AC, b, x, start + AF/nb; writes b32, x32 (+ reads AC32))
        cyc += 32.0 * host[0].n + 12.0 * host[0].refs() + 8.0 * host[0].n + (p.pre >= 1 ? 4.0 * host[0].n : 0.0);
        preDone = (L == 1 ? (!p.coarseDirect && p.coarseSweeps >= 1) : (p.pre >= 1)) ? 1 : 0;
    }
    cyc += cycleFrom(0, preDone);
    if (mixed) cyc += 20.0 * host[0].n;  // x += e (FP64 x read + write, FP32 e)
    return bytes + cycles * cyc;
}

// ---------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------
Solver::Solver(int device) : device_(device) {
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    AXS_HIP(hipHostMalloc((void**)&hctl_, sizeof(Ctl), hipHostMallocDefault));
    std::memset(hctl_, 0, sizeof(Ctl));
    axs_config_defaults(&cfg_);
}

// the learned guess's device state (defined here, before the destructor that
// deletes it)
struct Solver::NetSetup {
    NetGraphDev graph;
    NetWork work;
    Arena arena;
    std::vector<double*> vec;  // per level 1..netLevel_: restricted b, later the prolonged guess
    double* x = nullptr;       // the guess on netLevel_
};

Solver::~Solver() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (hctl_) (void)hipHostFree(hctl_);
    if (hhist_) (void)hipHostFree(hhist_);
    delete netSetup_;
    netSetup_ = nullptr;
    mg_.arena.release();
    stateArena_.release();
    meshArena_.release();
    bcArena_.release();
    colorArena_.release();
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Solver::setMesh(const axs_mesh_view& v) {
    if (running_) throw std::logic_error("solver is running");
    if (v.n_cells <= 0 || v.n_faces <= 0) throw std::invalid_argument("mesh has no cells or faces");
    const int N = v.n_cells, F = v.n_faces;
    if (!v.cell_face_start || v.cell_face_start[0] != 0) throw std::invalid_argument("cell_face_start[0] must be 0");
    for (int c = 0; c < N; c++)
        if (v.cell_face_start[c + 1] < v.cell_face_start[c]) throw std::invalid_argument("cell_face_start not monotone");
    const int S = v.cell_face_start[N];
    for (int k = 0; k < S; k++)
        if (v.cell_face_ids[k] < 0 || v.cell_face_ids[k] >= F) throw std::invalid_argument("cell_face_ids out of range");
    for (int f = 0; f < F; f++) {
        if (v.face_owner[f] < 0 || v.face_owner[f] >= N) throw std::invalid_argument("face_owner out of range");
        if (v.face_neighbor[f] < -1 || v.face_neighbor[f] >= N) throw std::invalid_argument("face_neighbor out of range");
    }
    HostMesh m;
    m.nCells = N;
    m.nFaces = F;
    m.cellCenterZ.assign(v.cell_center_z, v.cell_center_z + N);
    m.cellCenterR.assign(v.cell_center_r, v.cell_center_r + N);
    m.cellVolume.assign(v.cell_volume, v.cell_volume + N);
    m.cellArea2D.assign(v.cell_area2d, v.cell_area2d + N);
    m.cellFaceStart.assign(v.cell_face_start, v.cell_face_start + N + 1);
    m.cellFaceIDs.assign(v.cell_face_ids, v.cell_face_ids + S);
    m.faceOwner.assign(v.face_owner, v.face_owner + F);
    m.faceNeighbor.assign(v.face_neighbor, v.face_neighbor + F);
    m.faceGroup.assign(v.face_group, v.face_group + F);
    m.faceNormalZ.assign(v.face_normal_z, v.face_normal_z + F);
    m.faceNormalR.assign(v.face_normal_r, v.face_normal_r + F);
    m.faceCenterZ.assign(v.face_center_z, v.face_center_z + F);
    m.faceCenterR.assign(v.face_center_r, v.face_center_r + F);
    m.faceArea.assign(v.face_area, v.face_area + F);
    mesh_ = std::move(m);
    computeGeometry(mesh_, geom_);
    cellNeighbourCSR(mesh_, csrStart_, csrNb_);
    haveMesh_ = true;
    meshUploaded_ = false;
    mgBuilt_ = false;
    mixedAllocated_ = false;
    nColors_ = 0;
    colorStart_.clear();
    sweepBlocks_ = SweepBlocks{};
    nCases_ = 0;  // case ranges refer to the previous mesh
    casesReady_ = false;
}

void Solver::setBC(int var, const axs_bc_view& v) {
    if (running_) throw std::logic_error("solver is running");
    if (var < 0 || var >= AXS_VAR_COUNT) throw std::invalid_argument("bad variable");
    if (v.n_groups < 0 || (v.n_groups > 0 && (!v.type || !v.value))) throw std::invalid_argument("bad bc view");
    HostBC h;
    h.nGroups = v.n_groups;
    for (int g = 0; g < v.n_groups; g++) {
        const uint8_t t = v.type[g];
        if (t > AXS_BC_PULSATILE) throw std::invalid_argument("bad bc type");
        if ((t == AXS_BC_MICHAELIS_MENTEN || t == AXS_BC_HILL) && var < AXS_VAR_T)
            throw std::domain_error("reactive-wall kinetics apply to temperature and concentration only");
        if (t == AXS_BC_MICHAELIS_MENTEN || t == AXS_BC_HILL) h.kinetics = true;
        h.pulsatile.push_back(t == AXS_BC_PULSATILE);
        h.type.push_back(t == AXS_BC_PULSATILE ? (uint8_t)AXS_BC_DIRICHLET : t);  // memory_manger.cpp:96-100
        h.value.push_back(v.value[g]);
        h.length.push_back(v.length ? v.length[g] : 0.0);
        h.amplitude.push_back(v.amplitude ? v.amplitude[g] : 0.0);
        h.frequency.push_back(v.frequency ? v.frequency[g] : 0.0);
        // createBoundaryFieldHost defaults: rates 0, exponents 1 (memory_manger.cpp:70-82)
        h.vmax.push_back(v.vmax ? v.vmax[g] : 0.0);
        h.km.push_back(v.km ? v.km[g] : 0.0);
        h.hillN.push_back(v.hill_n ? v.hill_n[g] : 1.0);
        h.inhibM.push_back(v.inhib_m ? v.inhib_m[g] : 1.0);
        h.k2.push_back(v.inhib_k2 ? v.inhib_k2[g] : 0.0);
        h.v2.push_back(v.inhib_v2 ? v.inhib_v2[g] : 0.0);
        h.rtot.push_back(v.wall_resistance ? v.wall_resistance[g] : 0.0);
        // Km^n and K2^m once per group (createBoundaryFieldDevice, memory_manger.cpp:175-184)
        h.kmN.push_back(std::pow(h.km.back(), h.hillN.back()));
        h.k2M.push_back(std::pow(h.k2.back(), h.inhibM.back()));
    }
    bc_[var] = std::move(h);
}

// The p' multigrid settings of a configuration (ConfigMultigrid, solver_struct.h:167-190)
static MGParams mgParamsOf(const axs_config& c) {
    MGParams mp;
    mp.cycles = c.mg_max_iter;
    mp.weight = c.mg_weight;
    mp.pre = c.mg_pre_sweeps;
    mp.post = c.mg_post_sweeps;
    mp.coarseSweeps = c.mg_coarse_sweeps;
    mp.precision = c.precision;
    mp.coarseDirect = c.coarse_direct;
    mp.smoother = c.mg_smoother;
    mp.refine = c.mixed_refine;
    return mp;
}

void Solver::configure(const axs_config& c) {
    if (running_) throw std::logic_error("solver is running");
    if (c.precision != AXS_PREC_FP64 && c.precision != AXS_PREC_MIXED) throw std::invalid_argument("bad precision");
    if (c.mg_smoother != AXS_MG_SMOOTH_JACOBI && c.mg_smoother != AXS_MG_SMOOTH_RBGS)
        throw std::invalid_argument("bad multigrid smoother");
    if (c.mixed_refine != AXS_REFINE_CYCLE && c.mixed_refine != AXS_REFINE_SOLVE)
        throw std::invalid_argument("bad mixed-precision refinement");
    if (c.coarse_direct < 0) throw std::invalid_argument("coarse_direct must be 0 (off) or the refresh interval");
    cfg_ = c;
    haveCfg_ = true;
}

void Solver::uploadMesh() {
    meshArena_.release();
    const HostMesh& m = mesh_;
    DMesh d{};
    d.N = m.nCells;
    d.F = m.nFaces;
    d.S = (int)m.cellFaceIDs.size();
    d.cZ = meshArena_.upload(m.cellCenterZ);
    d.cR = meshArena_.upload(m.cellCenterR);
    d.vol = meshArena_.upload(m.cellVolume);
    d.invA2D = meshArena_.upload(geom_.invA2D);
    d.lsqInvZZ = meshArena_.upload(geom_.lsqInvZZ);
    d.lsqInvZR = meshArena_.upload(geom_.lsqInvZR);
    d.lsqInvRR = meshArena_.upload(geom_.lsqInvRR);
    d.faceStart = meshArena_.upload(m.cellFaceStart);
    d.faceIDs = meshArena_.upload(m.cellFaceIDs);
    d.slotNb = meshArena_.upload(csrNb_);
    d.maxSlots = 0;
    int minSlots = 1 << 30;
    for (int c = 0; c < m.nCells; c++) {
        const int k = m.cellFaceStart[c + 1] - m.cellFaceStart[c];
        d.maxSlots = std::max(d.maxSlots, k);
        minSlots = std::min(minSlots, k);
    }
    d.uniformSlots = (minSlots == d.maxSlots && m.cellFaceStart[0] == 0) ? d.maxSlots : 0;
    d.owner = meshArena_.upload(m.faceOwner);
    d.neighbor = meshArena_.upload(m.faceNeighbor);
    d.group = meshArena_.upload(m.faceGroup);
    d.nZ = meshArena_.upload(m.faceNormalZ);
    d.nR = meshArena_.upload(m.faceNormalR);
    d.fZ = meshArena_.upload(m.faceCenterZ);
    d.fR = meshArena_.upload(m.faceCenterR);
    d.area = meshArena_.upload(m.faceArea);
    d.invDPN = meshArena_.upload(geom_.invDPN);
    d.wP = meshArena_.upload(geom_.wP);
    d.dPB = meshArena_.upload(geom_.dPB);
    d.lsqWZ = meshArena_.upload(geom_.lsqWZ);
    d.lsqWR = meshArena_.upload(geom_.lsqWR);
    d.length2D = meshArena_.upload(geom_.length2D);
    {
        SlotGeometry sg;
        computeSlotGeometry(m, geom_, sg);
        d.sArea = meshArena_.upload(sg.area);
        d.sInvDPN = meshArena_.upload(sg.invDPN);
        d.sNZ = meshArena_.upload(sg.nz);
        d.sNR = meshArena_.upload(sg.nr);
        d.sW = meshArena_.upload(sg.w);
        d.sTZ = meshArena_.upload(sg.tz);
        d.sTR = meshArena_.upload(sg.tr);
        d.sLsqWZ = meshArena_.upload(sg.lsqWZ);
        d.sLsqWR = meshArena_.upload(sg.lsqWR);
        d.sLen = meshArena_.upload(sg.len2D);
        d.sFlags = meshArena_.upload(sg.flags);
    }
    dm_ = d;
    meshUploaded_ = true;
}

// allocateSimple / allocateCoefficients / allocateResiduals: zero initial state
// (memory_manger.cpp:270-285, 691-773, 802-832)
void Solver::allocateState() {
    stateArena_.release();
    const int N = dm_.N, F = dm_.F, S = dm_.S;
    DState s{};
    for (double** p : {&s.u, &s.v, &s.p, &s.pp, &s.uTmp, &s.vTmp, &s.ppTmp, &s.DU, &s.DV, &s.gPZ, &s.gPR, &s.gUZ,
                       &s.gUR, &s.gVZ, &s.gVR, &s.uOld, &s.vOld, &s.uOld2, &s.vOld2, &s.uAC, &s.ub, &s.uInv, &s.vAC,
                       &s.vb, &s.vInv, &s.ppAC, &s.ppb, &s.ppInv})
        *p = stateArena_.zeros<double>(N);
    s.mDot = stateArena_.zeros<double>(F);
    s.AFuv = stateArena_.zeros<double>(S);
    s.ppAF = stateArena_.zeros<double>(S);
    s.uvA = stateArena_.zeros<double2>(N);
    s.uvB = stateArena_.zeros<double2>(N);
    partialBlocks_ = blocksFor(N);
    s.partial = stateArena_.zeros<double>((size_t)partialBlocks_ * (9 + 8));
    s.ctl = stateArena_.zeros<Ctl>(1);
    s.hist = nullptr;
    histCap_ = 0;
    ds_ = s;
    // temperature / concentration (allocateSimple allocates the solved ones,
    // memory_manger.cpp:691-773)
    const bool want[2] = {cfg_.solve_energy != 0, cfg_.solve_concentration != 0};
    for (int i = 0; i < 2; i++) {
        sc_[i] = DScalar{};
        scAlloc_[i] = want[i];
        if (!want[i]) continue;
        DScalar& x = sc_[i];
        for (double** p : {&x.phi, &x.tmp, &x.old, &x.old2, &x.gZ, &x.gR, &x.AC, &x.b, &x.inv})
            *p = stateArena_.zeros<double>(N);
        x.AF = stateArena_.zeros<double>(S);
    }
    wall_ = DWall{};
    if (want[0] || want[1]) {
        wall_.cw = stateArena_.zeros<double>(F);
        wall_.ocr = stateArena_.zeros<double>(F);
    }
    std::memset(hctl_, 0, sizeof(Ctl));
    stateAllocated_ = true;
}

void Solver::allocateMixed() {
    if (!mixedAllocated_) {
        mixedArena_.release();
        const int N = dm_.N, S = dm_.S;
        MixedState x{};
        x.AF32 = mixedArena_.zeros<float>(S);
        for (float2** p : {&x.inv, &x.r, &x.d, &x.d2}) *p = mixedArena_.zeros<float2>(N);
        mx_ = x;
        mixedAllocated_ = true;
    }
    ds_.AF32 = mx_.AF32;
    ds_.uvInv32 = mx_.inv;
}

// createBoundaryFieldHost/Device (memory_manger.cpp:51-189) with the value of
// each pulsatile group at `time` (structs/boundary_struct.h:184-192)
void Solver::uploadBCs(double time) {
    const double twoPi = 2.0 * 3.14159265358979323846;
    auto values = [&](const HostBC& h) {
        std::vector<double> v(h.value);
        for (int g = 0; g < h.nGroups; g++)
            if (h.pulsatile[g]) v[g] = h.value[g] * (1.0 + h.amplitude[g] * std::sin(twoPi * h.frequency[g] * time));
        return v;
    };
    if (bcValueDev_[0] == nullptr) {
        bcArena_.release();
        for (int var = 0; var < AXS_VAR_COUNT; var++) {
            const HostBC& h = bc_[var];
            DBC d{};
            d.nGroups = h.nGroups;
            d.type = bcArena_.upload(h.type);
            bcValueDev_[var] = bcArena_.upload(values(h));
            d.value = bcValueDev_[var];
            d.length = bcArena_.upload(h.length);
            if (h.kinetics) {
                d.vmax = bcArena_.upload(h.vmax);
                d.km = bcArena_.upload(h.km);
                d.hillN = bcArena_.upload(h.hillN);
                d.inhibM = bcArena_.upload(h.inhibM);
                d.v2 = bcArena_.upload(h.v2);
                d.rtot = bcArena_.upload(h.rtot);
                d.kmN = bcArena_.upload(h.kmN);
                d.k2M = bcArena_.upload(h.k2M);
            }
            dbc_[var] = d;
        }
        // ppBC: the p types with every value zeroed (solver.cpp:788-795)
        ppbc_ = dbc_[AXS_VAR_P];
        ppbc_.value = bcArena_.zeros<double>(std::max(1, bc_[AXS_VAR_P].nGroups));
        return;
    }
    for (int var = 0; var < 2; var++) {
        const std::vector<double> v = values(bc_[var]);
        if (!v.empty())
            AXS_HIP(hipMemcpy(bcValueDev_[var], v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    }
}

Physics Solver::physics() const {
    const axs_config& c = cfg_;
    Physics ph{};
    ph.rho = c.rho;
    ph.mu = c.mu;
    ph.alphaU = c.alpha_u;
    ph.alphaP = c.alpha_p;
    ph.dt = c.transient ? c.dt : 0.0;
    ph.transient = c.transient ? 1 : 0;
    ph.conv = c.add_convection ? 1 : 0;
    // QUICK is clamped to upwind on load (file_manager.cpp:273-275) and falls
    // to upwind inside the kernel (solver_util.cuh:765-766)
    ph.scheme = (c.convection_scheme == AXS_CONV_QUICK) ? AXS_CONV_UPWIND : c.convection_scheme;
    ph.gradScheme = c.gradient_scheme;
    ph.nonOrtho = c.use_non_orth ? 1 : 0;
    return ph;
}

// One SIMPLE outer iteration, solver.cpp:1116-1273
void Solver::enqueueIteration(hipStream_t s) {
    const axs_config& c = cfg_;
    const Physics ph = physics();
    const bool needGrad = c.add_convection && (c.convection_scheme == AXS_CONV_SOU || c.convection_scheme == AXS_CONV_QUICK);
    const DBC &bu = dbc_[AXS_VAR_U], &bv = dbc_[AXS_VAR_V], &bp = dbc_[AXS_VAR_P];
    DState st = ds_;

    // the pressure gradient rides along with the velocity gradients when those
    // are needed (one walk over the geometry instead of two)
    const bool gradUV = ph.nonOrtho || needGrad;
    if (gradUV) launchGradUV(s, dm_, bu, bv, bp, st, ph.gradScheme, true);
    // mixed, fused FP32 sweeps: the momentum assembly also forms the FP64
    // defect b - A u0 the sweeps start from
    const std::vector<int> chunks = sweepChunks(c.linear_max_iter);
    bool fusedMixed = c.precision == AXS_PREC_MIXED && chunks.size() >= 2;
    for (int n : chunks) fusedMixed = fusedMixed && n >= 2;
    st.momR32 = fusedMixed ? mx_.r : nullptr;
    launchMomentum(s, dm_, bu, bv, bp, st, ph, gradUV);

    // momentum solves (LinearSolver::run, linear_solver.cu:91-148, 209-271)
    if (c.precision == AXS_PREC_MIXED) {
        // FP32 Jacobi on the FP64 defect: sweep 1 inside kMomDefect, sweeps
        // 2..L-1 in FP32, sweep L fused with u += delta
        const int L = c.linear_max_iter;
        float2 *d = mx_.d, *d2 = mx_.d2;
        if (fusedMixed) {
            // the same sweeps in the fused FP64 path's chunks: the first launch
            // starts from the defect kMomentum wrote, the last ends in u += delta
            for (size_t i = 0; i < chunks.size(); i++) {
                if (i == 0) launchMomSweep32BlockFromDefect(s, dm_, st, mx_, dsb_, chunks[i], d2);
                else if (i + 1 == chunks.size()) launchMomSweep32BlockApply(s, dm_, st, mx_, dsb_, chunks[i], d);
                else launchMomSweep32Block(s, dm_, st, mx_, dsb_, chunks[i], d, d2);
                std::swap(d, d2);
            }
        } else {
            launchMomDefect(s, dm_, st, mx_);
            // sweeps 2 .. L-1 in chunks of up to fuse_ per launch (the FP64 path's
            // blocks), avoiding a single sweep at the end where possible
            for (int left = L - 2; left > 0;) {
                int n = std::min(fuse_, left);
                if (n > 2 && left - n == 1) n--;
                if (n == 1) launchMomSweep32(s, dm_, st.ctl, mx_, d, d2);
                else launchMomSweep32Block(s, dm_, st, mx_, dsb_, n, d, d2);
                std::swap(d, d2);
                left -= n;
            }
            launchMomSweepApply(s, dm_, st, mx_, d, L >= 2 ? 1 : 0);
        }
    } else if (c.linear_solver == AXS_LINEAR_GS_RB && nColors_ > 0) {
        for (int it = 0; it < c.linear_max_iter; it++)
            for (int col = 0; col < nColors_; col++) {
                const int begin = colorStart_[col], count = colorStart_[col + 1] - begin;
                if (count > 0) launchGSColorUV(s, dm_, st, dOrder_, begin, count);
            }
    } else if (c.linear_max_iter >= 2) {
        // u, v -> (u,v) pairs -> ... -> (u,v) pairs -> u, v: the same sweeps, the
        // iterates in between interleaved; the last one lands in u, v directly
        // Fused: chunks of up to fuse_ sweeps per launch, the same sequence of
        // iterates; no launch reads and writes the same array.
        double2 *a = st.uvA, *b = st.uvB;
        for (size_t i = 0; i < chunks.size(); i++) {
            const bool first = i == 0, last = i + 1 == chunks.size();
            const double *uo = first ? st.u : nullptr, *vo = first ? st.v : nullptr;
            const double2* uvo = first ? nullptr : a;
            double *un = last ? st.u : nullptr, *vn = last ? st.v : nullptr;
            double2* uvn = last ? nullptr : (first ? a : b);
            if (chunks[i] == 1) launchJacobiUVPacked(s, dm_, st, uo, vo, uvo, un, vn, uvn);
            else launchJacobiUVBlock(s, dm_, st, dsb_, chunks[i], uo, vo, uvo, un, vn, uvn);
            if (!first) std::swap(a, b);
        }
    } else if (c.linear_max_iter == 1) {
        launchJacobiUV(s, dm_, st, st.u, st.v, st.uTmp, st.vTmp);
        launchCopy(s, st.u, st.uTmp, dm_.N, st.ctl);
        launchCopy(s, st.v, st.vTmp, dm_.N, st.ctl);
    }

    launchPPCoeff(s, dm_, bp, st, ph);
    launchFaceFlux(s, dm_, bu, bv, bp, st, ph);
    // p' = 0 before the solve (unless the learned guess supplies it): written
    // by the first right-hand-side launch
    const bool guess = c.use_multigrid && net_ != nullptr && netLevel_ > 0;

    MGParams mp = mgParamsOf(c);
    mp.l0Copies = c.precision == AXS_PREC_MIXED;
    mp.formInverse = formInverse_;
    for (int corr = 0; corr <= ph.nonOrtho; corr++) {
        const int cross = corr > 0 ? 1 : 0;
        if (cross) launchGradField(s, dm_, ppbc_, st.pp, st.gPZ, st.gPR, ph.gradScheme, st.ctl);
        // p' = 0 here: the right-hand side launch zeroes it (and writes b in
        // FP32 for the once-refined mixed solve)
        mp.startZero = corr == 0 && !guess;
        launchPPRhs(s, dm_, st, ph, cross, mp.startZero ? st.pp : nullptr);
        if (c.use_multigrid && corr == 0 && guess) {
            // operators, then the learned guess in place of p' = 0, then the cycles
            enqueueGuess(s, mp);
            mg_.enqueueSolve(s, mp, st.ctl, false);
        } else if (c.use_multigrid) {
            // the operator is the same for every corrector pass: build it once
            mg_.enqueueSolve(s, mp, st.ctl, corr == 0);
        } else {
            // ppSolver: the configured linear solver on p' (solver.cpp:921-924, 1203-1208)
            if (c.linear_solver == AXS_LINEAR_GS_RB && nColors_ > 0) {
                for (int it = 0; it < c.linear_max_iter; it++)
                    for (int col = 0; col < nColors_; col++) {
                        const int begin = colorStart_[col], count = colorStart_[col + 1] - begin;
                        if (count > 0)
                            launchGSColorSys(s, dm_, st.ppb, st.ppInv, st.ppAF, st.pp, dOrder_, begin, count, st.ctl);
                    }
            } else {
                double *x = st.pp, *t = st.ppTmp;
                for (int it = 0; it < c.linear_max_iter; it++) {
                    launchJacobiSys(s, dm_, st.ppb, st.ppInv, st.ppAF, x, t, st.ctl);
                    std::swap(x, t);
                }
                if (x != st.pp) launchCopy(s, st.pp, x, dm_.N, st.ctl);
            }
        }
    }
    launchGradPPUpdate(s, dm_, ppbc_, st, ph);
    launchMassFluxUpdate(s, dm_, bp, st, ph);
    if (scalarsOn()) enqueueScalars(s);
    if (nCases_ > 0) {
        // each case's residuals over blocks that start at its first cell, its
        // own convergence, then the cases done keep their fields
        DState sc = st;
        sc.partial = casePartial_;
        launchResidual(s, dm_, sc, caseBlocks_, dcases_.blockCells);
        if (scalarsOn()) launchResidualScalars(s, dm_, sc, sc_[0], sc_[1], caseBlocks_, dcases_.blockCells);
        launchFinishCases(s, sc, dcases_, caseBlocks_, scalarsOn());
        launchFreeze(s, dcases_, dm_.N, freezeCells_, dm_.F, freezeFaces_);
        return;
    }
    launchResidual(s, dm_, st, partialBlocks_);
    if (scalarsOn()) launchResidualScalars(s, dm_, st, sc_[0], sc_[1], partialBlocks_);
    launchFinish(s, st, partialBlocks_, scalarsOn());
}

void Solver::setCases(int count, const int* cellStart, const int* faceStart) {
    if (running_) throw std::logic_error("solver is running");
    if (!haveMesh_) throw std::logic_error("set the mesh before its cases");
    if (count < 0) throw std::invalid_argument("negative case count");
    caseCellStart_.clear();
    caseFaceStart_.clear();
    nCases_ = 0;
    casesReady_ = false;
    if (count == 0) return;
    if (!cellStart || !faceStart) throw std::invalid_argument("case ranges missing");
    if (cellStart[0] != 0 || faceStart[0] != 0 || cellStart[count] != mesh_.nCells || faceStart[count] != mesh_.nFaces)
        throw std::invalid_argument("case ranges must cover the mesh's cells and faces from 0");
    for (int c = 0; c < count; c++)
        if (cellStart[c + 1] <= cellStart[c] || faceStart[c + 1] <= faceStart[c])
            throw std::invalid_argument("every case needs cells and faces");
    // a case's faces connect only its own cells
    for (int c = 0; c < count; c++)
        for (int f = faceStart[c]; f < faceStart[c + 1]; f++) {
            const int o = mesh_.faceOwner[f], nb = mesh_.faceNeighbor[f];
            if (o < cellStart[c] || o >= cellStart[c + 1] || (nb >= 0 && (nb < cellStart[c] || nb >= cellStart[c + 1])))
                throw std::invalid_argument("case " + std::to_string(c) + " has a face outside its cells");
        }
    caseCellStart_.assign(cellStart, cellStart + count + 1);
    caseFaceStart_.assign(faceStart, faceStart + count + 1);
    nCases_ = count;
}

// Device tables of the cases (after the state is allocated: the snapshots
// cover the fields allocated for this run).
void Solver::prepareCases() {
    if (nCases_ == 0 || casesReady_) return;
    caseArena_.release();
    const int n = nCases_;
    std::vector<int> blockStart(n + 1, 0), cellCount(n), caseOfCell(mesh_.nCells), caseOfFace(mesh_.nFaces);
    std::vector<int2> blockCells;
    for (int c = 0; c < n; c++) {
        const int c0 = caseCellStart_[c], c1 = caseCellStart_[c + 1];
        cellCount[c] = c1 - c0;
        for (int b = c0; b < c1; b += kBlock) blockCells.push_back(make_int2(b, std::min(b + kBlock, c1)));
        blockStart[c + 1] = (int)blockCells.size();
        for (int i = c0; i < c1; i++) caseOfCell[i] = c;
        for (int f = caseFaceStart_[c]; f < caseFaceStart_[c + 1]; f++) caseOfFace[f] = c;
    }
    caseBlocks_ = (int)blockCells.size();
    DCases d{};
    d.n = n;
    d.blockStart = caseArena_.upload(blockStart);
    d.blockCells = caseArena_.upload(blockCells);
    d.cellCount = caseArena_.upload(cellCount);
    d.caseOfCell = caseArena_.upload(caseOfCell);
    d.caseOfFace = caseArena_.upload(caseOfFace);
    d.ctl = caseArena_.zeros<CaseCtl>(n);
    dcases_ = d;
    casePartial_ = caseArena_.zeros<double>((size_t)caseBlocks_ * (9 + 8));
    caseHost_.assign(n, CaseCtl{});
    // the fields a case done keeps (everything an iteration changes that a
    // later one or the output reads)
    FreezeSet fc{}, ff{};
    auto add = [&](FreezeSet& set, double* field, int len) {
        if (!field) return;
        set.field[set.n] = field;
        set.snap[set.n] = caseArena_.zeros<double>(len);
        set.n++;
    };
    for (double* f : {ds_.u, ds_.v, ds_.p, ds_.pp, sc_[0].phi, sc_[1].phi}) add(fc, f, mesh_.nCells);
    for (double* f : {ds_.mDot, wall_.cw, wall_.ocr}) add(ff, f, mesh_.nFaces);
    freezeCells_ = fc;
    freezeFaces_ = ff;
    casesReady_ = true;
}

// Every case runs again (a new run or time step): done / frozen / converged
// cleared; residual and scale values kept (the continuity scale is captured
// again at the first checks, as in a run of the case alone).
void Solver::resetCases() {
    if (nCases_ == 0) return;
    AXS_HIP(hipMemcpy(caseHost_.data(), dcases_.ctl, (size_t)nCases_ * sizeof(CaseCtl), hipMemcpyDeviceToHost));
    for (CaseCtl& k : caseHost_) {  // iterations keep counting over steps and continued runs
        k.done = 0;
        k.frozen = 0;
        k.converged = 0;
    }
    AXS_HIP(hipMemcpy(dcases_.ctl, caseHost_.data(), (size_t)nCases_ * sizeof(CaseCtl), hipMemcpyHostToDevice));
}

void Solver::caseStats(int c, int& iterations, int& converged, double* lastResidual, int& stopIteration) const {
    if (c < 0 || c >= nCases_) throw std::out_of_range("no such case");
    if (!casesReady_) throw std::logic_error("run the batch first");
    const CaseCtl& k = caseHost_[c];
    iterations = k.iters;
    converged = k.converged;
    stopIteration = k.stopIter;
    for (int i = 0; i < AXS_RES_COUNT; i++) lastResidual[i] = k.resVal[i];
}

// The scalar equation of field `which` (0 temperature, 1 concentration):
// diffusivity k / (rho cp) or D, the volumetric flux, under-relaxation with
// alpha_u for temperature only (solver.cpp:912, 993-1009)
ScalarPhysics Solver::scalarPhysics(int which) const {
    const Physics ph = physics();
    ScalarPhysics sp{};
    sp.gamma = which == 0 ? cfg_.k / (cfg_.rho * cfg_.cp) : cfg_.diffusivity;
    sp.fluxScale = 1.0 / cfg_.rho;
    sp.alpha = which == 0 ? cfg_.alpha_u : 1.0;
    sp.dt = ph.dt;
    sp.conv = ph.conv;
    sp.scheme = ph.scheme;
    sp.nonOrtho = ph.nonOrtho;
    sp.transient = ph.transient;
    return sp;
}

// Energy, then concentration, after the mass-flux update (solver.cpp:1222-1237)
void Solver::enqueueScalars(hipStream_t s) {
    const axs_config& c = cfg_;
    const bool needGrad = c.add_convection && (c.convection_scheme == AXS_CONV_SOU || c.convection_scheme == AXS_CONV_QUICK);
    for (int i = 0; i < 2; i++) {
        if (!scAlloc_[i]) continue;
        const DScalar& x = sc_[i];
        const DBC& bc = dbc_[AXS_VAR_T + i];
        const ScalarPhysics sp = scalarPhysics(i);
        if (sp.nonOrtho || needGrad) launchGradField(s, dm_, bc, x.phi, x.gZ, x.gR, c.gradient_scheme, ds_.ctl);
        launchScalar(s, dm_, bc, x, wall_, ds_.mDot, sp, ds_.ctl);
        // the field's linear solver (LinearSolver::run, linear_solver.cu:91-148)
        if (c.linear_solver == AXS_LINEAR_GS_RB && nColors_ > 0) {
            for (int it = 0; it < c.linear_max_iter; it++)
                for (int col = 0; col < nColors_; col++) {
                    const int begin = colorStart_[col], count = colorStart_[col + 1] - begin;
                    if (count > 0) launchGSColorSys(s, dm_, x.b, x.inv, x.AF, x.phi, dOrder_, begin, count, ds_.ctl);
                }
        } else {
            double *a = x.phi, *t = x.tmp;
            for (int it = 0; it < c.linear_max_iter; it++) {
                launchJacobiSys(s, dm_, x.b, x.inv, x.AF, a, t, ds_.ctl);
                std::swap(a, t);
            }
            if (a != x.phi) launchCopy(s, x.phi, a, dm_.N, ds_.ctl);
        }
    }
}

void Solver::dropGraph() {
    for (hipGraphExec_t* e : {&exec_, &execLag_})
        if (*e) {
            AXS_HIP(hipGraphExecDestroy(*e));
            *e = nullptr;
        }
    for (hipGraph_t* g : {&graph_, &graphLag_})
        if (*g) {
            AXS_HIP(hipGraphDestroy(*g));
            *g = nullptr;
        }
}

// One outer iteration captured as a graph (formInverse_ as given).
void Solver::captureIteration(hipGraph_t* graph, hipGraphExec_t* exec) {
    AXS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueueIteration(stream_);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();  // leave no capture error behind for the next call
        formInverse_ = true;
        throw;
    }
    AXS_HIP(hipStreamEndCapture(stream_, graph));
    AXS_HIP(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
    AXS_HIP(hipGraphUpload(*exec, stream_));
}

void Solver::buildGraph() {
    dropGraph();
    formInverse_ = true;
    captureIteration(&graph_, &exec_);
    size_t nodes = 0;
    AXS_HIP(hipGraphGetNodes(graph_, nullptr, &nodes));
    graphNodes_ = (int)nodes;
    if (cfg_.coarse_direct > 1 && cfg_.use_multigrid) {
        formInverse_ = false;
        captureIteration(&graphLag_, &execLag_);
        formInverse_ = true;
    }
    launches_ = 0;
}

// n outer iterations; with coarse_direct = K > 1 the first of every K forms
// the coarsest inverse and the others apply it.
void Solver::launchBatch(int n) {
    const int every = cfg_.coarse_direct > 1 ? cfg_.coarse_direct : 1;
    for (int i = 0; i < n; i++, launches_++)
        AXS_HIP(hipGraphLaunch(execLag_ && launches_ % every != 0 ? execLag_ : exec_, stream_));
}

// Read ctl and the new history entries, report them.
void Solver::pullControl(axs_progress_cb cb, void* user) {
    const int before = hctl_->nHist;
    AXS_HIP(hipMemcpyAsync(hctl_, ds_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
    AXS_HIP(hipStreamSynchronize(stream_));
    const int after = hctl_->nHist;
    if (after > before) {
        const int cnt = after - before;
        AXS_HIP(hipMemcpy(hhist_, ds_.hist + (size_t)before * kHistStride, (size_t)cnt * kHistStride * sizeof(double),
                          hipMemcpyDeviceToHost));
        static const char* names[AXS_RES_COUNT] = {"U", "V", "Continuity", "Temperature", "Concentration"};
        for (int i = 0; i < cnt; i++) {
            const double* r = hhist_ + (size_t)i * kHistStride;
            for (int j = 0; j < AXS_RES_COUNT; j++) history_[j].push_back(r[j]);
            if (cb) cb(user, (int)r[kHistStride - 1], AXS_RES_COUNT, names, r);
        }
    }
}

void Solver::run(axs_progress_cb cb, void* user, bool continueRun) {
    if (!haveMesh_) throw std::logic_error("no mesh");
    bool expected = false;
    if (!running_.compare_exchange_strong(expected, true)) throw std::logic_error("solver is already running");
    struct Guard {
        std::atomic<bool>& r;
        ~Guard() { r = false; }
    } guard{running_};
    AXS_HIP(hipSetDevice(device_));
    const auto t0 = std::chrono::steady_clock::now();
    axs_config& c = cfg_;
    // runCheck, solver.cpp:530-661
    if (c.max_iter < 1) c.max_iter = 50;
    if (c.check_conv < 1) c.check_conv = 1;
    if (c.linear_max_iter < 1) c.linear_max_iter = 20;
    if (c.structured_mesh) c.use_non_orth = 0;
    if (c.save_key_frame_iter < 1) c.save_key_frame_iter = 1;
    if (c.transient && (!std::isfinite(c.dt) || c.dt <= 0.0 || !std::isfinite(c.t_end) || c.t_end <= 0.0))
        throw std::invalid_argument("Transient run needs a positive time step and end time.");
    if (!std::isfinite(c.rho) || !std::isfinite(c.mu) || c.rho <= 0.0 || c.mu <= 0.0)
        throw std::invalid_argument("Solver needs positive density and viscosity.");
    // the fluid properties the scalars read (file_manager.cpp:287-291 rejects the same)
    if (c.solve_energy && (!std::isfinite(c.cp) || c.cp <= 0.0 || !std::isfinite(c.k) || c.k < 0.0))
        throw std::invalid_argument("Energy equation needs a positive heat capacity and a non-negative conductivity.");
    if (c.solve_concentration && (!std::isfinite(c.diffusivity) || c.diffusivity < 0.0))
        throw std::invalid_argument("Concentration equation needs a non-negative diffusion coefficient.");
    for (int v = 0; v < 2; v++)
        for (int g = 0; g < bc_[v].nGroups; g++)
            if (bc_[v].pulsatile[g]) {
                if (!c.transient) throw std::invalid_argument("Pulsatile inlet requires Transient Solver to be enabled.");
                if (!std::isfinite(bc_[v].value[g]) || !(bc_[v].amplitude[g] >= 0.0) || !(bc_[v].frequency[g] > 0.0))
                    throw std::invalid_argument("Pulsatile inlet needs a finite mean velocity, A >= 0, and f > 0 Hz.");
            }
    const bool mixed = c.precision == AXS_PREC_MIXED;
    if (mixed && c.linear_solver != AXS_LINEAR_JACOBI)
        throw std::domain_error("mixed precision momentum sweeps are Jacobi only");
    if (mixed && !c.use_multigrid) throw std::domain_error("mixed precision needs the multigrid pressure solve");

    if (!meshUploaded_) uploadMesh();
    std::string why;
    const bool cont = continueRun && canContinue(why);
    if (!cont) {
        allocateState();
        casesReady_ = false;  // the snapshots cover this run's fields
        currentIteration_ = 0;
        for (auto& h : history_) h.clear();
        contCells_ = mesh_.nCells;
        contFaces_ = mesh_.nFaces;
        contRefs_ = (int)mesh_.cellFaceIDs.size();
    }
    bcValueDev_[0] = nullptr;
    uploadBCs(0.0);
    if (!mgBuilt_ || mgBuiltLevels_ != c.mg_max_levels) {
        mg_.build(mesh_.nCells, csrStart_, csrNb_, c.mg_max_levels);
        mgBuilt_ = true;
        mgBuiltLevels_ = c.mg_max_levels;
    }
    mg_.bindLevel0(ds_.ppAC, ds_.ppb, ds_.ppAF, ds_.pp, ds_.ppTmp);
    if (c.use_multigrid && c.mg_smoother == AXS_MG_SMOOTH_RBGS) mg_.ensureColoring();  // no allocation in capture
    setupGuess();
    if (mixed) {
        allocateMixed();
        mg_.ensureMixed();
        if (c.coarse_direct && mg_.coarsestCells() > kDirectMax)
            throw std::domain_error("coarsest level too large for the dense FP64 solve");
        // level 0's FP32 operator straight from the p' assembly, and with one
        // defect per solve its right-hand side from the first one
        ds_.ppAC32 = mg_.dev[0].AC32;
        ds_.ppAF32 = mg_.dev[0].AF32;
        ds_.ppb32 = c.mixed_refine == AXS_REFINE_SOLVE ? mg_.dev[0].b32 : nullptr;
    } else {
        ds_.AF32 = nullptr;
        ds_.uvInv32 = nullptr;
        ds_.ppAC32 = ds_.ppAF32 = ds_.ppb32 = nullptr;
        if (c.coarse_direct && c.use_multigrid) {
            if (mg_.coarsestCells() > kDirectMax)
                throw std::domain_error("coarsest level too large for the dense FP64 solve");
            mg_.ensureInverse();
        }
    }
    prepareSweepBlocks();
    if (c.linear_solver == AXS_LINEAR_GS_RB && nColors_ == 0) {
        std::vector<int> order;
        nColors_ = greedyColoring(mesh_.nCells, csrStart_.data(), csrNb_.data(), colorStart_, order);
        colorArena_.release();
        dOrder_ = colorArena_.upload(order);
    }
    // residual history buffer for this run
    const int cap = c.max_iter + 1;
    if (cap > histCap_) {
        ds_.hist = stateArena_.zeros<double>((size_t)cap * kHistStride);
        histCap_ = cap;
    }
    if (cap > hhistCap_) {
        if (hhist_) AXS_HIP(hipHostFree(hhist_));
        AXS_HIP(hipHostMalloc((void**)&hhist_, (size_t)cap * kHistStride * sizeof(double), hipHostMallocDefault));
        hhistCap_ = cap;
    }
    prepareCases();
    resetCases();
    buildGraph();

    // control block (scaleVal persists across continued runs like the pinned
    // ConfigResidual::scaleVal does)
    AXS_HIP(hipMemcpy(hctl_, ds_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    Ctl& k = *hctl_;
    k.done = 0;
    k.iter = currentIteration_;
    k.k = 0;
    k.maxIter = c.max_iter;
    k.checkConv = c.check_conv;
    k.transient = c.transient ? 1 : 0;
    k.bdf2Step = 0;
    k.nHist = 0;
    k.histCap = histCap_;
    k.converged = 0;
    k.N = mesh_.nCells;
    for (int i = 0; i < AXS_RES_COUNT; i++) {
        k.enabled[i] = c.residual[i].enabled;
        k.norm[i] = c.residual[i].norm;
        k.scaling[i] = c.residual[i].scaling;
        k.tol[i] = c.residual[i].tol;
    }
    AXS_HIP(hipMemcpy(ds_.ctl, hctl_, sizeof(Ctl), hipMemcpyHostToDevice));
    stop_ = false;
    stopped_ = false;
    converged_ = false;
    const auto t1 = std::chrono::steady_clock::now();
    setupMs_ = std::chrono::duration<double, std::milli>(t1 - t0).count();

    auto iterateUntilDone = [&]() {
        int batch = 1;
        while (true) {
            if (stop_) {
                stopped_ = true;
                break;
            }
            const int remaining = hctl_->maxIter - hctl_->k;
            launchBatch(std::max(1, std::min(batch, remaining)));
            pullControl(cb, user);
            if (hctl_->done) break;
            batch = std::min(batch * 2, 64);
        }
    };
    // Frames are kept only by a continued transient run, which carries time
    // and the pulsatile phase on from the last frame (solver.cpp:890-900); a
    // transient run starting afresh first captures the state it starts from
    // (solver.cpp:1028-1031).
    if (!c.transient || !cont) frames_.clear();
    const double timeOffset = frames_.empty() ? 0.0 : frames_.back().time;
    if (!c.transient) {
        iterateUntilDone();
    } else {
        if (frames_.empty()) captureFrame(timeOffset);
        const int steps = (int)std::ceil(c.t_end / c.dt);
        const bool bdf2 = c.time_scheme == AXS_TIME_BDF2;
        for (int t = 0; t < steps && !stopped_; t++) {
            AXS_HIP(hipStreamSynchronize(stream_));
            // implicit at t_(n+1) (solver.cpp:1036-1056)
            uploadBCs(timeOffset + (double)(t + 1) * c.dt);
            launchStepBegin(stream_, dm_, ds_, bdf2 ? 1 : 0);
            for (int i = 0; i < 2; i++)
                if (scAlloc_[i]) launchRotate(stream_, dm_.N, sc_[i].old2, sc_[i].old, sc_[i].phi, bdf2 ? 1 : 0);
            AXS_HIP(hipMemcpyAsync(hctl_, ds_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
            AXS_HIP(hipStreamSynchronize(stream_));
            hctl_->done = 0;
            hctl_->k = 0;
            hctl_->converged = 0;
            hctl_->nHist = 0;
            hctl_->bdf2Step = (bdf2 && t > 0) ? 1 : 0;
            AXS_HIP(hipMemcpy(ds_.ctl, hctl_, sizeof(Ctl), hipMemcpyHostToDevice));
            resetCases();
            iterateUntilDone();
            // the keyframe interval and the last (or stopped) step, solver.cpp:1286-1294
            const bool lastStep = (t == steps - 1) || stopped_;
            if (t % c.save_key_frame_iter == 0 || lastStep) captureFrame(timeOffset + (double)(t + 1) * c.dt);
        }
    }
    currentIteration_ = hctl_->iter;
    converged_ = hctl_->converged != 0;
    if (nCases_ > 0)
        AXS_HIP(hipMemcpy(caseHost_.data(), dcases_.ctl, (size_t)nCases_ * sizeof(CaseCtl), hipMemcpyDeviceToHost));
    // output gradients with the final fields (solver.cpp:1330-1336)
    launchGradField(stream_, dm_, dbc_[AXS_VAR_U], ds_.u, ds_.gUZ, ds_.gUR, c.gradient_scheme, nullptr);
    launchGradField(stream_, dm_, dbc_[AXS_VAR_V], ds_.v, ds_.gVZ, ds_.gVR, c.gradient_scheme, nullptr);
    launchGradField(stream_, dm_, dbc_[AXS_VAR_P], ds_.p, ds_.gPZ, ds_.gPR, c.gradient_scheme, nullptr);
    for (int i = 0; i < 2; i++)  // solver.cpp:1338-1340
        if (scAlloc_[i])
            launchGradField(stream_, dm_, dbc_[AXS_VAR_T + i], sc_[i].phi, sc_[i].gZ, sc_[i].gR, c.gradient_scheme,
                            nullptr);
    AXS_HIP(hipStreamSynchronize(stream_));
    AXS_HIP(hipGetLastError());
    solveMs_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    ready_ = true;
}

// Benchmark: exactly n graph launches (outer iterations) from the current
// state with the convergence test disabled, timed with HIP events.
void Solver::runTimed(int n, double* ms) {
    if (!ready_ || !exec_) throw std::logic_error("run the solver once before timing it");
    if (n <= 0) throw std::invalid_argument("n must be positive");
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipMemcpy(hctl_, ds_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    Ctl saved = *hctl_;
    hctl_->done = 0;
    hctl_->k = 0;
    hctl_->maxIter = n;
    hctl_->converged = 0;
    hctl_->nHist = 0;
    for (int i = 0; i < AXS_RES_COUNT; i++) hctl_->enabled[i] = 0;
    AXS_HIP(hipMemcpy(ds_.ctl, hctl_, sizeof(Ctl), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    AXS_HIP(hipEventCreate(&e0));
    AXS_HIP(hipEventCreate(&e1));
    AXS_HIP(hipEventRecord(e0, stream_));
    launchBatch(n);
    AXS_HIP(hipEventRecord(e1, stream_));
    AXS_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    AXS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    AXS_HIP(hipMemcpy(hctl_, ds_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const int iter = hctl_->iter;
    *hctl_ = saved;
    hctl_->iter = iter;
    AXS_HIP(hipMemcpy(ds_.ctl, hctl_, sizeof(Ctl), hipMemcpyHostToDevice));
    currentIteration_ = iter;
}

// Time `repeats` eager launches of one kernel from the current state.  The
// launches run with no control block (they ignore ctl->done).  Every kernel
// but the face flux rewrites only buffers the next outer iteration overwrites
// (and from the same inputs, so with the same values); the face flux
// overwrites the mass flux, which is state carried between iterations, so it
// is saved before the launches and restored after them.
void Solver::benchKernel(int kernel, int repeats, double* ms, double* bytes, bool cold) {
    if (!ready_) throw std::logic_error("run the solver once before timing a kernel");
    if (repeats <= 0) throw std::invalid_argument("repeats must be positive");
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamSynchronize(stream_));
    DState st = ds_;
    st.ctl = nullptr;
    const Physics ph = physics();
    const double N = dm_.N, F = dm_.F, S = dm_.S;
    const bool lsq = cfg_.gradient_scheme == AXS_GRAD_LSQ;
    std::function<void()> launch;
    switch (kernel) {
    case AXS_KERNEL_JACOBI_UV:
        // the interleaved sweep (all but the first and last of each solve): reads
        // b, invAC, x of u and v + CSR + shared AF; writes x of u and v
        *bytes = 68.0 * N + 12.0 * S;
        launch = [&] { launchJacobiUVPacked(stream_, dm_, st, nullptr, nullptr, st.uvA, nullptr, nullptr, st.uvB); };
        break;
    case AXS_KERNEL_JACOBI_UV_FUSED:
        // fuse_ sweeps in one launch: the same distinct data as one sweep (the
        // halo rows and the block rows read again are not counted)
        if (fuse_ < 2) throw std::logic_error("fused momentum sweeps are not in use");
        *bytes = 68.0 * N + 12.0 * S;
        launch = [&] {
            launchJacobiUVBlock(stream_, dm_, st, dsb_, fuse_, nullptr, nullptr, st.uvA, nullptr, nullptr, st.uvB);
        };
        break;
    case AXS_KERNEL_MOMENTUM: {
        const bool needGrad = ph.nonOrtho || (cfg_.add_convection && ph.scheme == AXS_CONV_SOU);
        // cells: faceStart, vol, cR, u, v, p, the gradient inverse, 10 outputs (+
        // the u, v gradients); slots: slotNb, faceIDs, flags, area, 1/|d|, the
        // non-orthogonal vector, weight, mDot (gather), AF out (+ LSQ weights, or
        // the oriented normal and length)
        *bytes = 4.0 * N + 40.0 * N + (lsq ? 24.0 * N : 8.0 * N) + 80.0 * N + (needGrad ? 32.0 * N : 0.0) +
                 65.0 * S + (lsq ? 16.0 * S : 25.0 * S);
        launch = [&] { launchMomentum(stream_, dm_, dbc_[0], dbc_[1], dbc_[2], st, ph, false); };
        break;
    }
    case AXS_KERNEL_MG_SMOOTH: {
        if (mg_.dev.empty()) throw std::logic_error("no multigrid hierarchy");
        *bytes = 36.0 * N + 12.0 * S;
        DLevel L = mg_.dev[0];
        launch = [&, L] { launchMGJacobi(stream_, L, cfg_.mg_weight, nullptr); };
        break;
    }
    case AXS_KERNEL_FACE_FLUX:
        // faces: owner, neighbour, normal, area, 1/|d|, wP, mDot out (the
        // boundary faces' group and d_PB aside); cells: u, v, p, grad p, DU, DV
        *bytes = 56.0 * F + 56.0 * N;
        launch = [&] { launchFaceFlux(stream_, dm_, dbc_[0], dbc_[1], dbc_[2], st, ph); };
        break;
    case AXS_KERNEL_PP_COEFF:
        // cells: faceStart, DU, DV, AC and 1/AC out; slots: slotNb, area, 1/|d|,
        // the oriented normal, weight, AF out (boundary slots' face data aside)
        *bytes = 36.0 * N + 52.0 * S;
        launch = [&] { launchPPCoeff(stream_, dm_, dbc_[2], st, ph); };
        break;
    default: throw std::invalid_argument("unknown kernel id");
    }
    Arena saved;
    double* mDotSaved = nullptr;
    if (kernel == AXS_KERNEL_FACE_FLUX) {
        mDotSaved = saved.alloc<double>(dm_.F);
        AXS_HIP(hipMemcpyAsync(mDotSaved, ds_.mDot, (size_t)dm_.F * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    }
    launch();  // warm-up
    float f = 0.0f;
    if (!cold) {
        hipEvent_t e0, e1;
        AXS_HIP(hipEventCreate(&e0));
        AXS_HIP(hipEventCreate(&e1));
        AXS_HIP(hipEventRecord(e0, stream_));
        for (int i = 0; i < repeats; i++) launch();
        AXS_HIP(hipEventRecord(e1, stream_));
        AXS_HIP(hipEventSynchronize(e1));
        AXS_HIP(hipEventElapsedTime(&f, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        // cold caches: a kernel reads 512 MiB (more than the 4 MB L2 of each
        // XCD and the 256 MB last-level cache hold) before each launch, so
        // the launch finds its inputs in HBM only; the launch's time is what
        // `repeats` (flush, launch) pairs take beyond `repeats` flushes alone,
        // both back to back on the stream
        const size_t flushN = (size_t)64 << 20;  // doubles
        double* flush = saved.alloc<double>(flushN);
        double* sink = saved.alloc<double>(2048);
        launchFill(stream_, flush, 0.0, (int)flushN, nullptr);
        auto timed = [&](bool withKernel) {
            hipEvent_t e0, e1;
            AXS_HIP(hipEventCreate(&e0));
            AXS_HIP(hipEventCreate(&e1));
            launchStreamRead(stream_, flush, flushN, sink, 2048, -1.0);
            AXS_HIP(hipEventRecord(e0, stream_));
            for (int i = 0; i < repeats; i++) {
                launchStreamRead(stream_, flush, flushN, sink, 2048, -1.0);
                if (withKernel) launch();
            }
            AXS_HIP(hipEventRecord(e1, stream_));
            AXS_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            AXS_HIP(hipEventElapsedTime(&t, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return t;
        };
        const float flushOnly = timed(false);
        f = std::max(0.0f, timed(true) - flushOnly);
    }
    if (mDotSaved) {
        AXS_HIP(hipMemcpyAsync(ds_.mDot, mDotSaved, (size_t)dm_.F * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        AXS_HIP(hipStreamSynchronize(stream_));
    }
    AXS_HIP(hipGetLastError());
    *ms = (double)f / repeats;
}

// captureTimeFrame, solver.cpp:265-296: u, v, p copied to the host; the cap
// stops capture, not the solve.
void Solver::captureFrame(double time) {
    const int cap = cfg_.max_time_frames > 0 ? cfg_.max_time_frames : 500;
    if ((int)frames_.size() >= cap) return;
    AXS_HIP(hipStreamSynchronize(stream_));
    const size_t N = mesh_.nCells;
    Frame f;
    f.time = time;
    for (auto [dst, src] : {std::pair<std::vector<double>*, const double*>{&f.u, ds_.u}, {&f.v, ds_.v}, {&f.p, ds_.p},
                            {&f.t, sc_[0].phi}, {&f.c, sc_[1].phi}}) {
        if (!src) continue;  // a scalar not solved
        dst->resize(N);
        AXS_HIP(hipMemcpy(dst->data(), src, N * sizeof(double), hipMemcpyDeviceToHost));
    }
    frames_.push_back(std::move(f));
}

// wallConsumptionDiagnostic, solver_util.cu:480-531, on host copies of the
// concentration, the wall state and the mass flux (summed in face order).
void Solver::wallDiagnostic(double out[10]) {
    for (int i = 0; i < 10; i++) out[i] = 0.0;
    if (!scAlloc_[1]) throw std::logic_error("concentration is not solved");
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamSynchronize(stream_));
    const int N = mesh_.nCells, F = mesh_.nFaces;
    std::vector<double> conc(N), cw(F), ocr(F), md(F);
    AXS_HIP(hipMemcpy(conc.data(), sc_[1].phi, N * sizeof(double), hipMemcpyDeviceToHost));
    AXS_HIP(hipMemcpy(cw.data(), wall_.cw, F * sizeof(double), hipMemcpyDeviceToHost));
    AXS_HIP(hipMemcpy(ocr.data(), wall_.ocr, F * sizeof(double), hipMemcpyDeviceToHost));
    AXS_HIP(hipMemcpy(md.data(), ds_.mDot, F * sizeof(double), hipMemcpyDeviceToHost));
    const HostBC& bc = bc_[AXS_VAR_C];
    for (int f = 0; f < F; f++) {
        if (mesh_.faceNeighbor[f] >= 0) continue;
        const int ow = mesh_.faceOwner[f], g = mesh_.faceGroup[f];
        if (g < 0 || g >= bc.nGroups) continue;
        const uint8_t t = bc.type[g];
        if (t == AXS_BC_DIRICHLET && md[f] < 0.0) out[1] += -md[f] * bc.value[g];
        if (!(t == AXS_BC_MICHAELIS_MENTEN || t == AXS_BC_HILL)) continue;
        const double dPF = std::fabs((mesh_.faceCenterZ[f] - mesh_.cellCenterZ[ow]) * mesh_.faceNormalZ[f] +
                                     (mesh_.faceCenterR[f] - mesh_.cellCenterR[ow]) * mesh_.faceNormalR[f]);
        const double h = 1.0 / (dPF / cfg_.diffusivity + bc.rtot[g]);
        const double area = mesh_.faceArea[f], cp = conc[ow];
        out[0] += ocr[f];
        out[2] += area * h * cp;
        out[3] += 1.0;
        out[4] += cw[f];
        out[5] += dPF;
        out[6] += h;
        out[7] += cp;
        out[8] += bc.vmax[g];
        out[9] += bc.km[g];
    }
}

const Solver::Frame& Solver::frame(int index) const {
    if (index < 0 || index >= (int)frames_.size()) throw std::out_of_range("no such frame");
    return frames_[index];
}

bool Solver::canContinue(std::string& reason) const {  // solver.cpp:469-528
    if (running_ && !stateAllocated_) {
        reason = "Solver is already running.";
        return false;
    }
    if (!ready_ || !stateAllocated_) {
        reason = "Run the solver once before continuing.";
        return false;
    }
    if (mesh_.nCells != contCells_) {
        reason = "Mesh cell count changed.";
        return false;
    }
    if (mesh_.nFaces != contFaces_) {
        reason = "Mesh face count changed.";
        return false;
    }
    if ((int)mesh_.cellFaceIDs.size() != contRefs_) {
        reason = "Mesh connectivity changed.";
        return false;
    }
    if ((cfg_.solve_energy != 0) != scAlloc_[0]) {  // solver.cpp:486-504
        reason = "Temperature field selection changed.";
        return false;
    }
    if ((cfg_.solve_concentration != 0) != scAlloc_[1]) {
        reason = "Concentration field selection changed.";
        return false;
    }
    return true;
}

size_t Solver::fieldLength(int field) const {
    if (!haveMesh_) return 0;
    const size_t N = mesh_.nCells, F = mesh_.nFaces, S = mesh_.cellFaceIDs.size();
    if (field >= AXS_FIELD_U && field <= AXS_FIELD_CELL_RE) return N;
    if (field == AXS_FIELD_MDOT) return F;
    if (field == AXS_FIELD_DU || field == AXS_FIELD_DV) return N;
    if (field >= AXS_FIELD_U_AC && field <= AXS_FIELD_PP_AF)
        return ((field - AXS_FIELD_U_AC) % 4 == 3) ? S : N;
    if (field >= AXS_FIELD_U_OLD && field <= AXS_FIELD_GRAD_CR) return N;
    if (field >= AXS_FIELD_T_AC && field <= AXS_FIELD_C_OLD2) return (field == AXS_FIELD_T_AF || field == AXS_FIELD_C_AF) ? S : N;
    if (field == AXS_FIELD_WALL_CW || field == AXS_FIELD_WALL_OCR) return F;
    switch (field) {
    case AXS_GEOM_INV_A2D:
    case AXS_GEOM_LSQ_INV_ZZ:
    case AXS_GEOM_LSQ_INV_ZR:
    case AXS_GEOM_LSQ_INV_RR: return N;
    case AXS_GEOM_INV_DPN:
    case AXS_GEOM_WP:
    case AXS_GEOM_DPB:
    case AXS_GEOM_LSQ_WZ:
    case AXS_GEOM_LSQ_WR:
    case AXS_GEOM_LENGTH2D: return F;
    default: return 0;
    }
}

void Solver::getField(int field, double* out, size_t n) {
    const size_t len = fieldLength(field);
    if (len == 0) throw std::invalid_argument("unknown field");
    if (n != len) throw std::out_of_range("output length does not match the field");
    AXS_HIP(hipSetDevice(device_));
    auto down = [&](const double* src) {
        AXS_HIP(hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost));
    };
    auto host = [&](const std::vector<double>& v) { std::copy(v.begin(), v.end(), out); };
    const bool fieldNeedsState = field < AXS_GEOM_INV_DPN;
    if (fieldNeedsState && !stateAllocated_) throw std::logic_error("no solution yet");
    const int N = mesh_.nCells;
    switch (field) {
    case AXS_FIELD_U: down(ds_.u); break;
    case AXS_FIELD_V: down(ds_.v); break;
    case AXS_FIELD_P: down(ds_.p); break;
    case AXS_FIELD_PP: down(ds_.pp); break;
    case AXS_FIELD_GRAD_UZ: down(ds_.gUZ); break;
    case AXS_FIELD_GRAD_UR: down(ds_.gUR); break;
    case AXS_FIELD_GRAD_VZ: down(ds_.gVZ); break;
    case AXS_FIELD_GRAD_VR: down(ds_.gVR); break;
    case AXS_FIELD_GRAD_PZ: down(ds_.gPZ); break;
    case AXS_FIELD_GRAD_PR: down(ds_.gPR); break;
    case AXS_FIELD_MDOT: down(ds_.mDot); break;
    case AXS_FIELD_DU: down(ds_.DU); break;
    case AXS_FIELD_DV: down(ds_.DV); break;
    case AXS_FIELD_U_AC: down(ds_.uAC); break;
    case AXS_FIELD_U_B: down(ds_.ub); break;
    case AXS_FIELD_U_INVAC: down(ds_.uInv); break;
    case AXS_FIELD_U_AF: down(ds_.AFuv); break;
    case AXS_FIELD_V_AC: down(ds_.vAC); break;
    case AXS_FIELD_V_B: down(ds_.vb); break;
    case AXS_FIELD_V_INVAC: down(ds_.vInv); break;
    case AXS_FIELD_V_AF: down(ds_.AFuv); break;
    case AXS_FIELD_PP_AC: down(ds_.ppAC); break;
    case AXS_FIELD_PP_B: down(ds_.ppb); break;
    case AXS_FIELD_PP_INVAC: down(ds_.ppInv); break;
    case AXS_FIELD_PP_AF: down(ds_.ppAF); break;
    case AXS_FIELD_U_OLD: down(ds_.uOld); break;
    case AXS_FIELD_V_OLD: down(ds_.vOld); break;
    case AXS_FIELD_U_OLD2: down(ds_.uOld2); break;
    case AXS_FIELD_V_OLD2: down(ds_.vOld2); break;
    case AXS_FIELD_T: case AXS_FIELD_C: case AXS_FIELD_GRAD_TZ: case AXS_FIELD_GRAD_TR: case AXS_FIELD_GRAD_CZ:
    case AXS_FIELD_GRAD_CR: case AXS_FIELD_T_AC: case AXS_FIELD_T_B: case AXS_FIELD_T_INVAC: case AXS_FIELD_T_AF:
    case AXS_FIELD_C_AC: case AXS_FIELD_C_B: case AXS_FIELD_C_INVAC: case AXS_FIELD_C_AF: case AXS_FIELD_T_OLD:
    case AXS_FIELD_C_OLD: case AXS_FIELD_T_OLD2: case AXS_FIELD_C_OLD2: {
        const bool isT = field == AXS_FIELD_T || field == AXS_FIELD_GRAD_TZ || field == AXS_FIELD_GRAD_TR ||
                         (field >= AXS_FIELD_T_AC && field <= AXS_FIELD_T_AF) || field == AXS_FIELD_T_OLD ||
                         field == AXS_FIELD_T_OLD2;
        const DScalar& x = sc_[isT ? 0 : 1];
        if (!scAlloc_[isT ? 0 : 1]) throw std::logic_error(isT ? "temperature is not solved" : "concentration is not solved");
        const double* src = nullptr;
        switch (field) {
        case AXS_FIELD_T: case AXS_FIELD_C: src = x.phi; break;
        case AXS_FIELD_GRAD_TZ: case AXS_FIELD_GRAD_CZ: src = x.gZ; break;
        case AXS_FIELD_GRAD_TR: case AXS_FIELD_GRAD_CR: src = x.gR; break;
        case AXS_FIELD_T_AC: case AXS_FIELD_C_AC: src = x.AC; break;
        case AXS_FIELD_T_B: case AXS_FIELD_C_B: src = x.b; break;
        case AXS_FIELD_T_INVAC: case AXS_FIELD_C_INVAC: src = x.inv; break;
        case AXS_FIELD_T_AF: case AXS_FIELD_C_AF: src = x.AF; break;
        case AXS_FIELD_T_OLD: case AXS_FIELD_C_OLD: src = x.old; break;
        default: src = x.old2; break;
        }
        down(src);
        break;
    }
    case AXS_FIELD_WALL_CW:
    case AXS_FIELD_WALL_OCR:
        if (!wall_.cw) throw std::logic_error("no scalar is solved");
        down(field == AXS_FIELD_WALL_CW ? wall_.cw : wall_.ocr);
        break;
    case AXS_FIELD_CONTINUITY: {  // getMassImbalance, solver.cpp:298-331
        std::vector<double> md(mesh_.nFaces);
        AXS_HIP(hipMemcpy(md.data(), ds_.mDot, md.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int c = 0; c < N; c++) {
            double imb = 0.0;
            for (int k = mesh_.cellFaceStart[c]; k < mesh_.cellFaceStart[c + 1]; k++) {
                const int f = mesh_.cellFaceIDs[k];
                if (mesh_.faceOwner[f] == c) imb += md[f];
                else if (mesh_.faceNeighbor[f] == c) imb -= md[f];
            }
            out[c] = imb;
        }
        break;
    }
    case AXS_FIELD_VMAG:
    case AXS_FIELD_CELL_RE: {  // createDerivedVelocitySolutions, solver.cpp:221-263
        std::vector<double> u(N), v(N);
        AXS_HIP(hipMemcpy(u.data(), ds_.u, N * sizeof(double), hipMemcpyDeviceToHost));
        AXS_HIP(hipMemcpy(v.data(), ds_.v, N * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) {
            const double sp = std::sqrt(u[i] * u[i] + v[i] * v[i]);
            if (field == AXS_FIELD_VMAG) {
                out[i] = sp;
            } else {
                const double h = std::sqrt(std::max(mesh_.cellArea2D[i], 0.0));
                out[i] = (cfg_.mu > 0.0) ? cfg_.rho * sp * h / cfg_.mu : 0.0;
            }
        }
        break;
    }
    case AXS_GEOM_INV_DPN: host(geom_.invDPN); break;
    case AXS_GEOM_WP: host(geom_.wP); break;
    case AXS_GEOM_DPB: host(geom_.dPB); break;
    case AXS_GEOM_INV_A2D: host(geom_.invA2D); break;
    case AXS_GEOM_LSQ_WZ: host(geom_.lsqWZ); break;
    case AXS_GEOM_LSQ_WR: host(geom_.lsqWR); break;
    case AXS_GEOM_LSQ_INV_ZZ: host(geom_.lsqInvZZ); break;
    case AXS_GEOM_LSQ_INV_ZR: host(geom_.lsqInvZR); break;
    case AXS_GEOM_LSQ_INV_RR: host(geom_.lsqInvRR); break;
    case AXS_GEOM_LENGTH2D: host(geom_.length2D); break;
    default: throw std::invalid_argument("unknown field");
    }
}

int Solver::residualHistory(int res, double* out, int maxN) const {
    if (res < 0 || res >= AXS_RES_COUNT) throw std::invalid_argument("bad residual index");
    const auto& h = history_[res];
    const int n = std::min<int>(maxN, (int)h.size());
    if (out) std::copy(h.begin(), h.begin() + n, out);
    return (int)h.size();
}

// Algorithmic bytes of one outer iteration of this schedule (distinct
// elements read + written per launch; FP64 8 B, int32 4 B).
void Solver::setSweepFusion(int sweeps, int blockCells) {
    if (running_) throw std::logic_error("solver is running");
    if (sweeps < 1 || sweeps > kMaxFusedSweeps) throw std::invalid_argument("sweeps per launch must be 1 to 5");
    if (blockCells != 0 && blockCells != 512 && blockCells != 1024)
        throw std::invalid_argument("fused blocks hold 512 or 1024 cells (0: either)");
    fuseReq_ = sweeps;
    fuseCellsReq_ = blockCells;
    mgFuseL0_ = sweeps >= 2;
}

// Build (or reuse) the blocks and halo rings for the largest sweep count
// <= fuseReq_ (and, unless fuseCellsReq_ fixes it, the larger block size: its
// halo is the smaller share) whose halos fit the kernel's LDS and whose cells'
// rows fit its registers; fuse_ = 1 when none does.
void Solver::prepareSweepBlocks() {
    // one sweep per launch asked for (axs_set_sweep_fusion): the multigrid's
    // level 0 unfused too
    mg_.setFuseLevel0(mgFuseL0_);
    fuse_ = 1;
    for (int s = fuseReq_; s >= 2; s--)
        for (int cells : {1024, 512}) {
        if (fuseCellsReq_ != 0 && cells != fuseCellsReq_) continue;
        if (sweepBlocks_.cells != cells || sweepBlocks_.rings != s - 1) {
            // index ranges first (they keep each block's rows contiguous), else
            // breadth-first patches
            SweepBlocks sb;
            buildSweepBlocks(mesh_.nCells, csrStart_, csrNb_, cells, s - 1, kSweepRegSlots, sb);
            if (sb.maxSlots <= kSweepRegSlots && sb.maxHalo > sweepHaloMax(cells))
                buildSweepBlocks(mesh_.nCells, csrStart_, csrNb_, cells, s - 1, kSweepRegSlots, sb, true);
            if (sb.maxSlots > kSweepRegSlots || sb.maxHalo > sweepHaloMax(cells)) continue;
            sweepBlocks_ = std::move(sb);
            dropGraph();  // it holds the old blocks' addresses
            AXS_HIP(hipStreamSynchronize(stream_));
            sweepArena_.release();
            dsb_.cells = sweepBlocks_.cells;
            dsb_.rings = sweepBlocks_.rings;
            dsb_.nBlocks = sweepBlocks_.nBlocks;
            dsb_.haloStart = sweepArena_.upload(sweepBlocks_.haloStart);
            dsb_.ringEnd = sweepArena_.upload(sweepBlocks_.ringEnd);
            dsb_.halo = sweepArena_.upload(sweepBlocks_.halo.empty() ? std::vector<int>{0} : sweepBlocks_.halo);
            dsb_.own = sweepBlocks_.patches ? sweepArena_.upload(sweepBlocks_.own) : nullptr;
            dsb_.locOwn = sweepArena_.upload(sweepBlocks_.locOwn);
            dsb_.locHalo = sweepArena_.upload(sweepBlocks_.locHalo.empty() ? std::vector<int16_t>{-1}
                                                                           : sweepBlocks_.locHalo);
        }
        fuse_ = s;
        return;
        }
}

// Sweeps per launch for a momentum solve of `sweeps` Jacobi sweeps: chunks of
// fuse_, none of 1 when avoidable, and at least two launches (the first reads
// u, v and the last writes them).
std::vector<int> Solver::sweepChunks(int sweeps) const {
    std::vector<int> out;
    if (fuse_ >= 2) {
        for (int left = sweeps; left > 0;) {
            int n = std::min(fuse_, left);
            if (left - n == 1 && n > 2) n--;
            out.push_back(n);
            left -= n;
        }
    }
    if (out.size() < 2) out.assign(sweeps, 1);
    return out;
}

double Solver::bytesPerIteration() const {
    const double N = mesh_.nCells, F = mesh_.nFaces, S = (double)mesh_.cellFaceIDs.size();
    const axs_config& c = cfg_;
    const bool lsq = c.gradient_scheme == AXS_GRAD_LSQ;
    const bool needGrad = c.add_convection && (c.convection_scheme == AXS_CONV_SOU || c.convection_scheme == AXS_CONV_QUICK);
    // Distinct bytes each kernel reads and writes with this layout: per-cell
    // arrays, the slot-ordered face data of the cell loops (per slot: slotNb 4,
    // the oriented geometry it uses, faceIDs 4 + flags 1 where it reads mDot),
    // face arrays in the per-face kernels.
    const double gradGeom = lsq ? (28.0 * N + 20.0 * S) : (12.0 * N + 37.0 * S);
    double b = 0.0;
    const bool gradUV = c.use_non_orth || needGrad;
    if (gradUV) b += gradGeom + 3 * 8.0 * N + 6 * 8.0 * N;                                 // grad u, v, p (one pass)
    b += (gradUV ? 16.0 * N : gradGeom + 8.0 * N)                                          // grad p read, or computed
         + 65.0 * S + 8.0 * 4 * N + (gradUV ? 32.0 * N : 0.0)                                // slots, cells
         + 8.0 * (gradUV ? 8 : 10) * N;                                                     // outputs
    const bool fused = c.linear_solver == AXS_LINEAR_JACOBI && c.precision != AXS_PREC_MIXED;
    b += (fused ? sweepChunks(c.linear_max_iter).size() : c.linear_max_iter) * (68.0 * N + 12.0 * S);  // u, v sweeps
    b += 16.0 * N + 44.0 * S + 8.0 * S + 16.0 * N;                                          // pp coeff
    b += 8.0 * 6 * N + 8.0 * 9 * F + 8.0 * F;                                               // face flux
    b += 8.0 * N;                                                                           // pp = 0
    const int passes = c.use_non_orth ? 2 : 1;
    for (int p = 0; p < passes; p++) {
        b += 13.0 * S + 4.0 * N + 8.0 * N;                                                   // pp rhs
        if (p > 0) b += gradGeom + 8.0 * N + 16.0 * N + 44.0 * S + 32.0 * N;                 // grad p', cross
        b += mg_.host.empty() ? 0.0 : mg_.bytesPerSolve(mgParamsOf(c), p == 0);
    }
    b += gradGeom + 8.0 * N + 16.0 * N + 8.0 * 6 * N + 8.0 * 3 * N;                         // grad p' + update
    b += 8.0 * 10 * F + 8.0 * 4 * N + 8.0 * F;                                              // mass flux
    b += 8.0 * 6 * N + 12.0 * S + 13.0 * S;                                                 // residuals
    return b;
}

// ---------------------------------------------------------------------------
// autotuner (build extension): workgroup size per kernel family and level
// ---------------------------------------------------------------------------
void Solver::autotune(int repeats, double* msDefault, double* msTuned) {
    if (running_) throw std::logic_error("solver is running");
    if (!stateAllocated_ || !meshUploaded_) throw std::logic_error("run the solver before autotuning");
    if (repeats <= 0) throw std::invalid_argument("repeats must be positive");
    AXS_HIP(hipSetDevice(device_));
    AXS_HIP(hipStreamSynchronize(stream_));
    hipEvent_t e0, e1;
    AXS_HIP(hipEventCreate(&e0));
    AXS_HIP(hipEventCreate(&e1));
    auto timeIt = [&](const std::function<void()>& launch) {
        launch();
        AXS_HIP(hipEventRecord(e0, stream_));
        for (int r = 0; r < repeats; r++) launch();
        AXS_HIP(hipEventRecord(e1, stream_));
        AXS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        AXS_HIP(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / repeats;
    };
    const int cands[] = {64, 128, 256};
    double def = 0.0, best = 0.0;
    // pick the fastest candidate of one family; set(c) applies a size
    auto tune = [&](const std::function<void(int)>& set, const std::function<void()>& launch) {
        double bt = 1e300;
        int bc = kBlock;
        for (int c : cands) {
            set(c);
            const double t = timeIt(launch);
            if (c == kBlock) def += t;
            if (t < bt) {
                bt = t;
                bc = c;
            }
        }
        set(bc);
        best += bt;
    };
    DState st = ds_;
    st.ctl = nullptr;
    // the interleaved momentum sweep, writing the sweep scratch buffers
    tune([&](int c) { dm_.sweepBlock = c; },
         [&] { launchJacobiUVPacked(stream_, dm_, st, nullptr, nullptr, st.uvA, nullptr, nullptr, st.uvB); });
    // sweeps fused per launch and the fused block size, timed over a whole FP64
    // momentum solve (scratch in, scratch out); the default is 3 sweeps of
    // 512-cell blocks
    if (cfg_.linear_solver == AXS_LINEAR_JACOBI && cfg_.precision != AXS_PREC_MIXED && cfg_.linear_max_iter >= 2) {
        const int L = cfg_.linear_max_iter;
        auto solveTime = [&] {
            const std::vector<int> ch = sweepChunks(L);
            return timeIt([&] {
                for (int n : ch) {
                    if (n == 1) launchJacobiUVPacked(stream_, dm_, st, nullptr, nullptr, st.uvA, nullptr, nullptr, st.uvB);
                    else launchJacobiUVBlock(stream_, dm_, st, dsb_, n, nullptr, nullptr, st.uvA, nullptr, nullptr, st.uvB);
                }
            });
        };
        // the default request first (whatever count its halos allow), then
        // every count and block size that fits
        const int options[][2] = {{kMaxFusedSweeps, 0}, {1, 0},    {2, 512}, {2, 1024}, {3, 512},
                                  {3, 1024},             {4, 512}, {4, 1024}, {5, 512}, {5, 1024}};
        double bt = 1e300;
        int bs = 1, bcells = 0;
        for (size_t i = 0; i < sizeof(options) / sizeof(options[0]); i++) {
            fuseReq_ = options[i][0];
            fuseCellsReq_ = options[i][1];
            prepareSweepBlocks();
            if (i > 0 && fuse_ != options[i][0]) continue;  // halo or rows do not fit
            const double t = solveTime();
            if (i == 0) def += t;
            if (t < bt) {
                bt = t;
                bs = fuse_;
                bcells = options[i][1];
            }
        }
        fuseReq_ = bs;
        fuseCellsReq_ = bcells;
        prepareSweepBlocks();
        best += bt;
    }
    // each multigrid level outside the single-workgroup tail: smoothing,
    // residual and prolongation kernels (they write only the level's scratch)
    if (cfg_.use_multigrid && mgBuilt_) {
        const bool fp32 = cfg_.precision == AXS_PREC_MIXED;
        const int L = (int)mg_.host.size();
        const int tail = mg_.tailLevel(fp32);
        const double w = cfg_.mg_weight;
        for (int l = 0; l < L; l++) {
            if (tail >= 0 && l >= tail) break;
            DLevel& d = mg_.dev[l];
            const bool coarsest = mg_.host[l].coarsest();
            tune([&](int c) { d.block = c; },
                 [&] {
                     if (fp32) {
                         launchMGJacobi32(stream_, d, (float)w, nullptr);
                         if (!coarsest) {
                             launchMGResidual32(stream_, d, nullptr);
                             launchMGProlongSmooth32(stream_, d, mg_.dev[l + 1], (float)w, nullptr);
                         }
                     } else {
                         launchMGJacobi(stream_, d, w, nullptr);
                         if (!coarsest) {
                             launchMGResidual(stream_, d, nullptr);
                             launchMGProlongSmooth(stream_, d, mg_.dev[l + 1], w, nullptr);
                         }
                     }
                 });
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    AXS_HIP(hipGetLastError());
    if (msDefault) *msDefault = def;
    if (msTuned) *msTuned = best;
}

// ---------------------------------------------------------------------------
// learned initial guess (build extension)
// ---------------------------------------------------------------------------
void Solver::setLearnedGuess(Mpnn* net, int level) {
    if (running_) throw std::logic_error("solver is running");
    if (level < -1 || level == 0) throw std::invalid_argument("guess level must be -1 or a coarse level >= 1");
    net_ = net;
    netLevelReq_ = level;
    netLevel_ = -1;
}

// Allocates everything the guess needs before the iteration graph is captured.
void Solver::setupGuess() {
    netLevel_ = -1;
    delete netSetup_;
    netSetup_ = nullptr;
    if (!net_ || !cfg_.use_multigrid) return;
    const int L = (int)mg_.host.size();
    int k = -1;
    if (netLevelReq_ >= 1) {
        if (netLevelReq_ >= L) throw std::out_of_range("guess level beyond the multigrid hierarchy");
        k = netLevelReq_;
    } else {
        for (int l = 1; l < L; l++)
            if (mg_.host[l].n <= kGuessCells) {
                k = l;
                break;
            }
    }
    if (k < 1) return;  // a single level: nothing coarser to run on
    netSetup_ = new NetSetup();
    netSetup_->graph.build(mg_.host[k].n, mg_.host[k].start, mg_.host[k].nb);
    netSetup_->work.ensure(netSetup_->graph.g.nPad, net_->layers());
    netSetup_->vec.assign(k + 1, nullptr);
    for (int l = 1; l <= k; l++) netSetup_->vec[l] = netSetup_->arena.zeros<double>(mg_.host[l].n);
    netSetup_->x = netSetup_->arena.zeros<double>(mg_.host[k].n);
    netLevel_ = k;
}

// b restricted to level k (sums over aggregates), the net's guess there, and
// that guess prolonged (piecewise constant) to p' on level 0.
void Solver::enqueueGuess(hipStream_t s, const MGParams& mp) {
    const int k = netLevel_;
    const Ctl* ctl = ds_.ctl;
    NetSetup& ns = *netSetup_;
    mg_.enqueueOperators(s, mp, ctl);
    const double* in = ds_.ppb;
    for (int l = 0; l < k; l++) {
        const DLevel& d = mg_.dev[l];
        netLaunchRestrict(s, mg_.host[l + 1].n, d.memberStart, d.members, in, ns.vec[l + 1], ctl);
        in = ns.vec[l + 1];
    }
    net_->enqueueForward(s, ns.graph.g, mg_.dev[k].AC, mg_.dev[k].AF, ns.vec[k], ns.x, ns.work);
    const double* src = ns.x;
    for (int l = k - 1; l >= 0; l--) {
        double* out = l == 0 ? ds_.pp : ns.vec[l];
        netLaunchProlong(s, mg_.host[l].n, mg_.dev[l].toCoarse, src, out, ctl);
        src = out;
    }
}

void Solver::stats(axs_stats& s) const {
    std::memset(&s, 0, sizeof(s));
    s.iterations = currentIteration_;
    s.converged = converged_ ? 1 : 0;
    s.stopped = stopped_ ? 1 : 0;
    s.n_cells = mesh_.nCells;
    s.n_faces = mesh_.nFaces;
    s.n_face_refs = (int)mesh_.cellFaceIDs.size();
    s.n_colors = nColors_;
    s.mg_levels = (int)mg_.host.size();
    for (int l = 0; l < (int)mg_.host.size() && l < 32; l++) s.mg_level_cells[l] = mg_.host[l].n;
    s.setup_ms = setupMs_;
    s.solve_ms = solveMs_;
    for (int i = 0; i < AXS_RES_COUNT; i++) s.last_residual[i] = history_[i].empty() ? 0.0 : history_[i].back();
    s.bytes_per_iteration = haveMesh_ ? bytesPerIteration() : 0.0;
    s.vcycle_bytes = mg_.host.empty() ? 0.0 : mg_.bytesPerSolve(mgParamsOf(cfg_), true);
    s.sweeps_per_launch = fuse_;
    s.graph_nodes = graphNodes_;
    s.sweep_block_cells = fuse_ >= 2 ? sweepBlocks_.cells : 0;
}

}  // namespace axs
