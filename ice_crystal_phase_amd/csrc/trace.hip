// Fused trace + bin kernels for gfx950.
//
// One persistent launch covers every record of a batch. The host cuts each
// record into tiles of consecutive ray ids; workgroups pull tiles from an
// atomic queue until the queue is empty (the exit condition every wave
// reaches), so uneven records and uneven ray lengths balance across the 256
// CUs. Each lane traces whole rays and adds every emission to the record's
// histogram as a u64 fixed-point quantum (rint(power * 2^Q)): integer sums are
// associative, so the histogram is bitwise independent of tile order, wave
// scheduling and atomic arrival order. Drop waves merge the emissions of a
// round that fall in one cell (adjacent lanes) before the HBM atomic; hex
// workgroups add into an LDS cell table flushed once per tile. Pole quanta and
// audit counters stay in lane registers and are reduced per wave before one
// atomic per counter.
//
// The record's u edges are staged in LDS when they fit (nt <= kLdsEdges-1) so
// the row lookup (phase_bin.hpp) reads LDS, not L2.
#include "kernels.hpp"

namespace icephase {

constexpr int kLdsEdges = 2048;

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct LaneCounters {
    unsigned long long fwd, bwd, samples, invalid, snapped, escaped, unresolved, hits;
    uint32_t status[9];
    uint32_t atomics;  // cell atomicAdds issued (per tile, so 32 bits per lane suffice)
};

// Counts the drop tile keeps per wave rather than per lane: events of a round
// counted where the whole wave runs (popcounts of ballots), so they take no
// lane registers through the tracing loops.
struct WaveTally {
    unsigned long long samples, atomics;
    uint32_t status[9];
};

__device__ inline void flush_counters(LaneCounters& c, unsigned long long* stats, const WaveTally& w)
{
    const unsigned long long v[9] = {c.fwd,     c.bwd,     c.samples,    c.invalid,         c.snapped,
                                     c.escaped, c.unresolved, c.hits, (unsigned long long)c.atomics};
    const unsigned long long wv[9] = {0, 0, w.samples, 0, 0, 0, 0, 0, w.atomics};
    const uint32_t slot[9] = {kStatForward, kStatBackward, kStatSamples,    kStatInvalid,    kStatSnapped,
                              kStatEscaped, kStatUnresolved, kStatInternalHits, kStatCellAtomics};
    const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const unsigned long long s = wave_sum(v[k]) + wv[k];
        if (lead && s) atomicAdd(&stats[slot[k]], s);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const unsigned long long s = wave_sum((unsigned long long)c.status[k]) + w.status[k];
        if (lead && s) atomicAdd(&stats[kStatStatus0 + k], s);
    }
}

// Status s of every lane where `on`, tallied for the wave (the whole wave calls it).
__device__ inline void tally_status(WaveTally& w, bool on, uint32_t s)
{
    s = s <= 8 ? s : 7;
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) w.status[k] += (uint32_t)__popcll(__ballot(on && s == k));
}

// Predicated adds keep the status array in registers (no dynamic indexing).
__device__ inline void count_status(LaneCounters& c, uint32_t s)
{
    s = s <= 8 ? s : 7;
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) c.status[k] += (s == k) ? 1u : 0u;
}

// Workgroup-private cell table of the hex kernels. A hex prism at one
// orientation has flat faces, so parallel rays that follow the same face
// sequence leave in exactly the same direction: a record's emissions fall in a
// few dozen cells (the 2^28-ray column record: 9.6 emissions per ray in 60
// cells, the hottest taking 12 % of them), and u64 atomicAdds into HBM would
// serialize on those cells. Each workgroup instead adds its quanta into an LDS
// open-addressing table (key = cell + 1, 0 = empty; slots claimed by atomicCAS,
// sums by LDS u64 atomicAdd) and adds the table into the record's histogram
// once per tile. A cell that finds no slot within kCellProbes goes straight to
// HBM. Integer sums are associative, so the histogram is the same bits either
// way. The DIRECT instantiation adds every quantum straight to HBM (the
// measured alternative, and the overflow path for every cell).
constexpr uint32_t kCellSlots = 2048;  // a power of two
constexpr uint32_t kCellProbes = 16;

template <bool DIRECT = false>
struct CellTable {
    uint32_t* keys;
    unsigned long long* sums;

    __device__ void add(unsigned long long* bins, uint32_t cell, unsigned long long q, LaneCounters& c) const
    {
        if constexpr (DIRECT) {
            atomicAdd(&bins[cell], q);
            ++c.atomics;
            return;
        }
        const uint32_t key = cell + 1u;
        uint32_t h = (cell * 2654435761u) >> 21;  // top 11 bits: kCellSlots slots
        for (uint32_t p = 0; p < kCellProbes; ++p) {
            uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (k == 0u) k = atomicCAS(&keys[h], 0u, key);  // the old key: 0 when this lane claimed it
            if (k == 0u || k == key) {
                atomicAdd(&sums[h], q);
                return;
            }
            h = (h + 1u) & (kCellSlots - 1u);
        }
        atomicAdd(&bins[cell], q);
        ++c.atomics;
    }
    // Every thread of the workgroup: zero the table (kernel start).
    __device__ void clear() const
    {
        for (uint32_t s = threadIdx.x; s < kCellSlots; s += blockDim.x) {
            keys[s] = 0u;
            sums[s] = 0ull;
        }
    }
    // Every thread of the workgroup, between barriers: add the table into the
    // histogram and clear it.
    __device__ void flush(unsigned long long* bins, LaneCounters& c) const
    {
        for (uint32_t s = threadIdx.x; s < kCellSlots; s += blockDim.x) {
            const uint32_t k = keys[s];
            if (k) {
                atomicAdd(&bins[k - 1u], sums[s]);
                ++c.atomics;
                keys[s] = 0u;
                sums[s] = 0ull;
            }
        }
    }
};

// Largest of v over the wave, merged into a non-negative FP64 counter by its
// bit pattern (non-negative doubles order like their u64 bits).
__device__ inline void wave_max_into(double v, unsigned long long* slot)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = gt_max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// A hex emission whose cell the FP32 decision leaves open: its FP64 direction
// (after the node rotation) and quantum, binned later by the FP64 rules.
struct HexDeferred {
    double x, y, z;
    unsigned long long q;
};
constexpr uint32_t kHexRing = 256;  // deferred hex emissions per wave (a power of two, at least 2 x 64)

// phase_bin's FP64 rules for a deferred hex emission (its unit checks already
// passed). Not inlined: the hex tile calls it between rounds of the splitting
// chain, whose live registers the FP64 binning would otherwise add to.
template <class EdgeLoad>
__device__ __attribute__((noinline)) CellHit bin_deferred_hex(const BinFrame& g, uint32_t nt, uint32_t np, double dx,
                                                              double dy, double dz, EdgeLoad edge)
{
    const d3 d = D3(dx, dy, dz);
    const double x = dot(g.e0, d), y = dot(g.e1, d), z = dot(g.ki, d);
    const double nrm = sqrt(x * x + y * y + z * z);
    return bin_exact(nt, np, x, y, z, nrm, edge);
}

// Cell bookkeeping of a binned hex emission: pole quanta and audit counters in
// lane registers, cell quanta into the workgroup's cell table.
template <class Table>
__device__ inline void record_hex(const DevRecord& R, const CellHit& h, unsigned long long q, const Table& table,
                                  LaneCounters& c)
{
    if (h.kind == 2) {
        ++c.invalid;
        return;
    }
    c.snapped += h.snapped;
    if (h.kind == 1)
        c.fwd += q;
    else if (h.kind == -1)
        c.bwd += q;
    else if (q)
        table.add(R.bins, h.index, q, c);
}

// Hex tile. Each lane traces whole rays; the wave runs its 64 rays' chains in
// lockstep, one interface per round (hex::enter, then hex::encounter up to
// max_hits times), so that every round's emissions are binned where the whole
// wave runs. phase_bin of an emission (phase_math.hpp:37-81 after the node
// rotation, phase_math.hpp:10): the unit checks on the spot, then the FP32
// decision (bin_fast_approx), which bins ~96 % of them; the rest go to the
// wave's LDS ring and are binned by the FP64 rules 64 at a time, between rays
// (inside a ray's rounds only when the ring could not take another round).
// The FP64 binning then runs with every lane busy and outside the splitting
// chain, whose registers it would otherwise add to. Integer quanta make the histogram
// independent of when an emission is binned. AREA: ray i enters at R.area[i]
// and carries the weight qf relative to the record's quantum (atlas nodes).
template <bool AREA, class EdgeLoad, class Table>
__device__ inline void hex_tile(const DevRecord& R, uint64_t first, uint64_t count, HexDeferred* ring, EdgeLoad edge,
                                const Table& table, LaneCounters& c)
{
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    double ray_balance = 0, face_balance = 0;
    unsigned long long rq = 0;
    // ring positions: changed only where every lane of the wave runs
    uint32_t head = 0, tail = 0;
    auto drain = [&](uint32_t n) {
        __builtin_amdgcn_wave_barrier();
        if (lane < n) {
            const HexDeferred e = ring[(head + lane) & (kHexRing - 1)];
            record_hex(R, bin_deferred_hex(R.frame, R.nt, R.np, e.x, e.y, e.z, edge), e.q, table, c);
        }
        head += n;
    };
    // One emission per lane where `sink`; called by the whole wave.
    auto emit = [&](bool sink, d3 d, double power, double qf) {
        bool defer = false;
        unsigned long long q = 0;
        if (sink) {
            if constexpr (AREA)
                q = quantize(power * qf * R.qfactor, R.qbits);
            else
                q = quantize(power * R.qfactor, R.qbits);
            rq += q;
            ++c.samples;
            if (R.rotated) {
                // body -> ensemble (hex_phase/include/ice_crystal/phase_math.hpp:10, same
                // operation order: column images summed left to right)
                const double* m = R.rot;
                d = D3((m[0] * d.x + m[1] * d.y) + m[2] * d.z, (m[3] * d.x + m[4] * d.y) + m[5] * d.z,
                       (m[6] * d.x + m[7] * d.y) + m[8] * d.z);
            }
            bool valid = is_unit64(d);
            if (valid) {
                const double x = dot(R.frame.e0, d), y = dot(R.frame.e1, d), z = dot(R.frame.ki, d);
                const double nrm = sqrt(x * x + y * y + z * z);
                valid = nrm > 0 && finite64(nrm) && fabs(nrm - 1) <= 512 * DBL_EPSILON;
            }
            if (!valid) {
                ++c.invalid;
            } else {
                const CellHit h = bin_fast_approx(R.nt, R.np, R.frame, d);
                if (h.kind == 0) {
                    if (q) table.add(R.bins, h.index, q, c);
                } else {
                    defer = true;
                }
            }
        }
        const unsigned long long m = __ballot(defer);
        if (m) {
            if (defer) ring[(tail + (uint32_t)__popcll(m & below)) & (kHexRing - 1)] = HexDeferred{d.x, d.y, d.z, q};
            tail += (uint32_t)__popcll(m);
            if (tail - head > kHexRing - 64) drain(64);  // so that the next round fits
        }
    };
    for (uint64_t base = 0; base < count; base += blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        const bool have = i < count;
        hex::Chain ch;
        ch.alive = false;
        ch.sink = false;
        double qf = 1.0;
        rq = 0;
        if (have) {
            hex::Entry e;
            if constexpr (AREA) {
                const AreaRay a = R.area[first + i];
                e = hex::Entry{a.pos, a.face, true};
                qf = a.qf;
            } else {
                e = hex::sample_entry(R.hex, first + i);
            }
            hex::enter(ch, R.hex, e);
        }
        emit(ch.sink, ch.dir, ch.pw, qf);  // exterior reflection
        for (uint32_t k = 0; k < R.hex.max_hits; ++k) {
            if (__ballot(ch.alive) == 0) break;
            ch.sink = false;
            if (ch.alive) hex::encounter(ch, R.hex);
            emit(ch.sink, ch.dir, ch.pw, qf);  // transmission of this encounter
        }
        if (ch.alive) hex::finish(ch, R.hex);
        while (tail - head >= 64) drain(64);
        if (have) {
            c.escaped += rq;
            if constexpr (AREA) {
                c.unresolved += quantize(ch.out.unresolved * qf * R.qfactor, R.qbits);
                ray_balance = gt_max(ray_balance, fabs(ch.out.escaped + ch.out.unresolved - 1));
                face_balance = gt_max(face_balance, ch.out.max_balance);
            } else {
                c.unresolved += quantize(ch.out.unresolved * R.qfactor, R.qbits);
            }
            c.hits += ch.out.internal_hits;
            count_status(c, ch.out.status);
        }
    }
    if (tail != head) drain(tail - head);  // the tile's remaining deferred emissions
    if constexpr (AREA) {
        wave_max_into(ray_balance, &R.stats[kStatRayBalance]);
        wave_max_into(face_balance, &R.stats[kStatFaceBalance]);
    }
}

// Cell bookkeeping of one binned emission: pole quanta and audit counters stay
// in lane registers; a cell gets the quantum by one u64 atomicAdd.
__device__ inline void record_hit(const DevRecord& R, const CellHit& h, unsigned long long q, LaneCounters& c)
{
    if (h.kind == 2) {
        ++c.invalid;
        return;
    }
    c.snapped += h.snapped;
    if (h.kind == 1)
        c.fwd += q;
    else if (h.kind == -1)
        c.bwd += q;
    else if (q) {
        atomicAdd(&R.bins[h.index], q);
        ++c.atomics;
    }
}

// An emission whose cell the FP32 decision (bin_fast) leaves open: its FP32
// direction and power, binned later by the FP64 rules.
struct Deferred {
    float x, y, z, pw;
};
constexpr uint32_t kRing = 128;  // deferred emissions per wave (a power of two, at least 2 x 64)

// phase_bin of a deferred drop emission: bin_raw's checks and FP64 rules (the
// FP32 decision already failed, and where it succeeds it agrees with them).
template <class EdgeLoad>
__device__ inline void bin_deferred(const DevRecord& R, const Deferred& e, EdgeLoad edge, LaneCounters& c)
{
    const unsigned long long q = quantize((double)e.pw, R.qbits);
    const d3 d = D3((double)e.x, (double)e.y, (double)e.z);
    CellHit h{0, 2, 0};
    if (finite64(d)) {
        const double x = dot(R.frame.e0, d), y = dot(R.frame.e1, d), z = dot(R.frame.ki, d);
        const double n = sqrt(x * x + y * y + z * z);
        if (n > 0 && finite64(n)) h = bin_exact(R.nt, R.np, x, y, z, n, edge);
    }
    record_hit(R, h, q, c);
}

// Drop tile. Lanes first acquire a ray that hits the drop: ids come from the
// tile's LDS counter (one wave-aggregated atomic per round), rays outside the
// bounding sphere are counted as misses on the spot, and the entry bracket of a
// candidate is set up here (foot of the perpendicular, else a secant search for
// the minimum of f along the line). The acquired rays are then traced together,
// one encounter per round of the whole wave: with no tail tolerance every
// hitting ray runs the same number of encounters, so the wave stays converged
// apart from Newton iteration counts. The sequence of operations applied to each
// ray is the per-ray algorithm the CPU mirror runs.
//
// Emissions are binned where the whole wave runs (after the entry and after each
// encounter round). The FP32 decision bins ~96 % of them on the spot; the rest
// go to the wave's LDS ring and are binned by the FP64 rules 64 at a time, so the
// FP64 path runs with every lane busy instead of in nearly every round for a
// lane or two (with 4 % undecided, 93 % of 64-lane rounds hold one). Integer
// quanta make the histogram independent of when an emission is binned.
template <class EdgeLoad>
__device__ inline void drop_tile(const DevRecord& R, uint64_t first, uint64_t count, uint32_t* s_next,
                                 Deferred* ring, EdgeLoad edge, LaneCounters& c, WaveTally& w)
{
    using namespace drop;
    const Config& C = R.drop;
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const float rmax2 = C.shape.sphere ? 1.0f : C.shape.rmax * C.shape.rmax;
    bool exhausted = false;
    unsigned long long rq = 0;
    // ring positions: changed only where every lane of the wave runs, so every
    // lane holds the same values
    uint32_t head = 0, tail = 0;
    auto drain = [&](uint32_t n) {
        __builtin_amdgcn_wave_barrier();
        if (lane < n) bin_deferred(R, ring[(head + lane) & (kRing - 1)], edge, c);
        head += n;
    };
    // One emission per lane where `valid`; called by the whole wave. Adjacent
    // lanes trace adjacent rays, so their emissions of one round often share a
    // cell (the headline oblate-drop batch: 63 emissions per round fall in 25
    // cells, always as runs of adjacent lanes): a segmented sum over each run of
    // equal cells leaves one u64 atomicAdd per run, issued by its first lane.
    auto emit = [&](bool valid, const Ray& r) {
        bool defer = false, counted = false;
        float pw = 0.0f;
        uint32_t cell = 0xffffffffu;  // no cell: no quantum to add
        unsigned long long sum = 0;
        if (valid) {
            pw = r.j.power();
            if (pw > 0.0f) {
                const unsigned long long q = quantize((double)pw, R.qbits);
                rq += q;
                counted = true;
                const d3 d = D3((double)r.dir.x, (double)r.dir.y, (double)r.dir.z);
                if (finite64(d)) {
                    const CellHit h = bin_fast_approx(R.nt, R.np, R.frame, d);
                    if (h.kind == 0) {
                        cell = h.index;
                        sum = q;
                    } else {
                        defer = true;
                    }
                } else {
                    defer = true;
                }
            }
        }
        {
            // every lane shuffles: a lane outside the exec mask would hand an
            // undefined value to its neighbour
            const uint32_t prev = __shfl_up(cell, 1, 64);
            const bool first_of_run = lane == 0 || prev != cell;
            // run number of each lane: first-of-run lanes at or below it
            const uint32_t run = (uint32_t)__popcll(__ballot(first_of_run) & (below | (1ull << lane)));
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {  // sum of the lane's run from the lane on
                const unsigned long long v = __shfl_down(sum, off, 64);
                const uint32_t vrun = __shfl_down(run, off, 64);  // every lane shuffles
                if (lane + off < 64 && vrun == run) sum += v;
            }
            const bool add = first_of_run && cell != 0xffffffffu && sum;
            if (add) atomicAdd(&R.bins[cell], sum);
            w.atomics += (uint32_t)__popcll(__ballot(add));
            w.samples += (uint32_t)__popcll(__ballot(counted));
        }
        const unsigned long long m = __ballot(defer);
        if (m) {
            if (defer) ring[(tail + (uint32_t)__popcll(m & below)) & (kRing - 1)] = Deferred{r.dir.x, r.dir.y, r.dir.z, pw};
            tail += (uint32_t)__popcll(m);
            if (tail - head >= 64) drain(64);  // leaves fewer than 64, so the next round fits
        }
    };
    for (;;) {
        bool have = false;
        float u = 0.0f, v = 0.0f, rho2 = 0.0f, ta = 0.0f, tm = 0.0f;
        f3 o = F3(0.0f, 0.0f, 0.0f);
        bool need = !exhausted;
        bool missed = false;  // this lane's id of the last round missed the drop
        for (;;) {
            tally_status(w, missed, kMiss);
            missed = false;
            const unsigned long long m = __ballot(need);
            if (m == 0) break;
            const uint32_t n = (uint32_t)__popcll(m);
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(s_next, n);
            base = __shfl(base, leader, 64);
            if (!need) continue;
            const uint64_t k = (uint64_t)base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
            if (k >= count) {
                exhausted = true;
                need = false;
                continue;
            }
            const uint64_t id = first + k;
            const uint32_t ix = (uint32_t)(id % C.w), iy = (uint32_t)(id / C.w);
            u = ((float)(2u * ix + 1u) / (float)C.w - 1.0f) * C.h;
            v = ((float)(2u * iy + 1u) / (float)C.hgt - 1.0f) * C.h;
            rho2 = fmaf(v, v, u * u);
            if (!(rho2 < rmax2)) {
                missed = true;
                continue;
            }
            o = axpy(C.e1, v, C.e0 * u);
            if (C.shape.sphere) {
                have = true;
                need = false;
                continue;
            }
            ta = -sqrtf(rmax2 - rho2);
            tm = 0.0f;
            float g0;
            const float fm = surface_slope(C.shape, o, C.ki, g0);
            if (fm < 0.0f) {
                have = true;
                need = false;
                continue;
            }
            // foot outside: secant minimization of f along the line from the foot
            float t0 = 0.0f, t1 = -g0 * sqrtf(rho2);
            for (int it = 0; it < 8; ++it) {
                t1 = t1 < ta ? ta : t1;
                t1 = t1 > -ta ? -ta : t1;
                float g1;
                const float f1 = surface_slope(C.shape, axpy(C.ki, t1, o), C.ki, g1);
                if (f1 < 0.0f) {
                    tm = t1;
                    have = true;
                    break;
                }
                if (g1 == g0) break;
                const float t2 = t1 - g1 * (t1 - t0) / (g1 - g0);
                t0 = t1;
                g0 = g1;
                if (fabsf(t2 - t1) <= 0x1p-20f) break;
                t1 = t2;
            }
            if (have)
                need = false;
            else
                missed = true;
        }
        if (__ballot(have) == 0) break;
        // From here every lane runs each round; lanes without a ray only take
        // part in the emissions.
        Ray ray;
        Split sp;
        sp.ok = false;
        sp.has_t = false;
        f3 p = F3(0.0f, 0.0f, 0.0f);
        uint32_t status = kLimit, hits = 0;
        float unresolved = 0.0f;
        bool alive = false, inner = false;
        if (have) {
            float t;
            if (C.shape.sphere) {
                t = -sqrtf(1.0f - rho2);
            } else {
                const float rm = 1.0f + C.shape.c[0];
                const float g2 = fmaf(rm, rm, -rho2);
                const float x0 = g2 > 0.0f ? -sqrtf(g2) : fmaf(tm - ta, 0.5f, ta);
                t = newton(C.shape, o, C.ki, ta, tm, true, x0);
            }
            p = axpy(C.ki, t, o);
            ray.dir = C.ki;
            ray.e0 = C.e0;
            ray.e1 = C.e1;
            ray.j = Jones{c32{1.0f, 0.0f}, c32{0.0f, 0.0f}, c32{0.0f, 0.0f}, c32{1.0f, 0.0f}};
            sp = split(ray, outward_normal(C.shape, p), C.n_ext, C.n_int);
            if (!sp.ok) {
                status = kNumeric;
                unresolved = 1.0f;
            } else if (!sp.has_t) {
                status = kExhausted;
            } else {
                ray = sp.t;
                alive = true;
                inner = true;
            }
        }
        emit(have && sp.ok, sp.r);  // external reflection
        for (uint32_t hit = 0; hit < C.max_hits; ++hit) {
            if (__ballot(alive) == 0) break;
            bool out = false;
            if (alive) {
                const float pw = ray.j.power();
                if (pw == 0.0f) {
                    status = kExhausted;
                    alive = false;
                } else if (pw <= C.tail) {
                    status = kTail;
                    alive = false;
                } else {
                    const float b = dot(p, ray.dir);
                    float tx = 0.0f;
                    bool found = true;
                    if (C.shape.sphere) {
                        const float cc = dot(p, p) - 1.0f;
                        float disc = fmaf(b, b, -cc);
                        if (disc < 0.0f) disc = 0.0f;
                        tx = -b + sqrtf(disc);
                    } else {
                        const float cb = dot(p, p) - rmax2;
                        float disc = fmaf(b, b, -cb);
                        if (disc < 0.0f) disc = 0.0f;
                        const float tb = -b + sqrtf(disc);
                        const float ts = 0x1p-16f;
                        if (!(ts < tb) || !(surface(C.shape, axpy(ray.dir, ts, p)) < 0.0f))
                            found = false;
                        else
                            tx = newton(C.shape, p, ray.dir, ts, tb, false, -2.0f * b);
                    }
                    if (!found) {
                        status = kUnresolved;
                        alive = false;
                    } else {
                        const f3 q = axpy(ray.dir, tx, p);
                        const f3 nout = outward_normal(C.shape, q);
                        sp = split(ray, nout * -1.0f, C.n_int, C.n_ext);
                        if (!sp.ok) {
                            status = kNumeric;
                            alive = false;
                        } else {
                            ++hits;
                            out = sp.has_t;
                            ray = sp.r;
                            p = q;
                        }
                    }
                }
            }
            emit(out, sp.t);  // transmitted branch of this encounter
        }
        if (inner) {
            unresolved = ray.j.power();
            if (status == kLimit && unresolved == 0.0f) status = kExhausted;
        }
        if (have) {
            c.escaped += rq;
            c.unresolved += quantize((double)unresolved, R.qbits);
            c.hits += hits;
        }
        tally_status(w, have, status);
        rq = 0;
    }
    if (tail != head) drain(tail - head);  // the tile's remaining deferred emissions
}

// Persistent tile loop of one record kind. LDS_EDGES: the record's u edges are
// staged in LDS (nt < kLdsEdges); otherwise they are read from global memory.
template <int KIND, bool LDS_EDGES, bool AREA = false, bool DIRECT = false>
__device__ inline void trace_loop(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    __shared__ double s_edges[LDS_EDGES ? kLdsEdges : 1];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cached;
    __shared__ uint32_t s_next;
    // per wave: the ring of deferred drop emissions (256 threads = 4 waves)
    __shared__ Deferred s_ring[KIND == kKindDrop ? 4 * kRing : 1];
    // hex: the workgroup's cell table (flushed into the record's histogram after every tile)
    __shared__ uint32_t s_keys[KIND == kKindHex ? kCellSlots : 1];
    __shared__ unsigned long long s_sums[KIND == kKindHex ? kCellSlots : 1];
    // hex: per wave, the ring of deferred emissions (256 threads = 4 waves)
    __shared__ HexDeferred s_hring[KIND == kKindHex ? 4 * kHexRing : 1];
    const CellTable<DIRECT> table{s_keys, s_sums};
    if constexpr (KIND == kKindHex && !DIRECT) table.clear();
    if (threadIdx.x == 0) s_cached = 0xffffffffu;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            s_tile = atomicAdd(queue, 1u);
            s_next = 0u;
        }
        __syncthreads();
        const uint64_t t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= map.total) break;
        // record of tile t: last k with tile_start[k] <= t (uniform scalar search)
        uint32_t lo = 0, hi = map.n;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (map.tile_start[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const uint32_t rec = __builtin_amdgcn_readfirstlane(map.recs_of_kind[lo]);
        const DevRecord& R = recs[rec];
        const uint64_t first = (t - map.tile_start[lo]) * R.tile_rays;
        const uint64_t count = R.rays - first < R.tile_rays ? R.rays - first : R.tile_rays;
        LaneCounters c{};
        WaveTally w{};
        if constexpr (LDS_EDGES) {
            if (s_cached != rec) {
                __syncthreads();
                for (uint32_t k = threadIdx.x; k <= R.nt; k += blockDim.x) s_edges[k] = R.edges[k];
                __syncthreads();
                if (threadIdx.x == 0) s_cached = rec;
                __syncthreads();
            }
            if constexpr (KIND == kKindHex)
                hex_tile<AREA>(R, first, count, s_hring + (threadIdx.x >> 6) * kHexRing, [&](int k) { return s_edges[k]; }, table, c);
            else
                drop_tile(R, first, count, &s_next, s_ring + (threadIdx.x >> 6) * kRing, [&](int k) { return s_edges[k]; }, c,
                          w);
        } else {
            const double* e = R.edges;
            if constexpr (KIND == kKindHex)
                hex_tile<AREA>(R, first, count, s_hring + (threadIdx.x >> 6) * kHexRing, [e](int k) { return e[k]; }, table, c);
            else
                drop_tile(R, first, count, &s_next, s_ring + (threadIdx.x >> 6) * kRing, [e](int k) { return e[k]; }, c,
                          w);
        }
        if constexpr (KIND == kKindHex && !DIRECT) {
            __syncthreads();  // every lane's quanta are in the table
            table.flush(R.bins, c);
        }
        flush_counters(c, R.stats, w);
    }
}

// Waves per SIMD the drop kernel is compiled for. Uncapped it runs 2 waves per
// SIMD; the serial Newton chains and the cell atomics' latency want more waves
// than the register file holds without spills. Measured on the first 48 records
// of the oblate workload at this build: 3 waves (168 VGPRs, 220 B/lane of
// spills) 504 ms, 4 waves (128 VGPRs, 396 B/lane) 472 ms, 5 waves (96 VGPRs,
// 528 B/lane) 716 ms; an earlier build ran 2 waves (uncapped) in 668 ms.
#ifndef ICEPHASE_DROP_WAVES
#define ICEPHASE_DROP_WAVES 4
#endif
// Waves per SIMD the hex kernels are compiled for: two, with the splitting
// chain's registers (the incoming ray and both outgoing rays, 17 FP64 values
// each) and a few spills.
#ifndef ICEPHASE_HEX_WAVES
#define ICEPHASE_HEX_WAVES 2
#endif
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_DROP_WAVES, 8)))
    icephase_trace_drop(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindDrop, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, true>(recs, map, queue);
}
// hex records whose rays enter at the points of an area quadrature (atlas nodes)
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex_area(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, true, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex_area_wide(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, false, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_DROP_WAVES, 8)))
    icephase_trace_drop_wide(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindDrop, false>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex_wide(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, false>(recs, map, queue);
}
// the hex kernels with every quantum added straight to HBM (no cell table)
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex_direct(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, true, false, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ICEPHASE_HEX_WAVES, 8)))
    icephase_trace_hex_wide_direct(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, false, false, true>(recs, map, queue);
}

// Bin caller directions (C-ABI iph_bin_directions).
extern "C" __global__ void __launch_bounds__(256)
    icephase_bin_dirs(BinFrame g, uint32_t nt, uint32_t np, const double* __restrict__ edges, uint64_t n,
                      const double* __restrict__ dirs, uint32_t* __restrict__ index, int32_t* __restrict__ kind)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const CellHit h = bin_direction(g, nt, np, D3(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]),
                                        [&](int e) { return edges[e]; });
        index[k] = h.index;
        kind[k] = h.kind;
    }
}

// Zero every record's histogram and counters, and the tile queue. A kernel, not
// hipMemsetAsync, so that the whole batch can live in a captured graph.
extern "C" __global__ void __launch_bounds__(256)
    icephase_zero(const DevRecord* __restrict__ recs, uint32_t nrec, uint32_t* __restrict__ queue)
{
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) queue[threadIdx.x] = 0u;
    for (uint32_t r = blockIdx.y; r < nrec; r += gridDim.y) {
        const DevRecord& R = recs[r];
        if (!R.owns) continue;
        const uint64_t n = (uint64_t)R.nt * R.np;
        for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
            R.bins[k] = 0ull;
        if (blockIdx.x == 0)
            for (uint32_t k = threadIdx.x; k < kStatCount; k += blockDim.x) R.stats[k] = 0ull;
    }
}

}  // namespace icephase

namespace icephase {

hipError_t launch_zero(const DevRecord* d_recs, uint32_t nrec, uint32_t* d_queue, hipStream_t s)
{
    hipLaunchKernelGGL(icephase_zero, dim3(256, nrec < 65535 ? nrec : 65535), dim3(256), 0, s, d_recs, nrec, d_queue);
    return hipGetLastError();
}

// kind: kKindDrop or kKindHex; wide: edges too large for LDS; direct: hex
// quanta straight to HBM (no workgroup cell table).
hipError_t launch_trace(int kind, bool wide, const DevRecord* d_recs, TileMap map, uint32_t* d_queue, uint32_t blocks,
                        uint32_t threads, hipStream_t s, bool direct)
{
    if (kind == kKindHex && direct) {
        if (wide)
            hipLaunchKernelGGL(icephase_trace_hex_wide_direct, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
        else
            hipLaunchKernelGGL(icephase_trace_hex_direct, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    } else if (kind == kKindHex) {
        if (wide)
            hipLaunchKernelGGL(icephase_trace_hex_wide, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
        else
            hipLaunchKernelGGL(icephase_trace_hex, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    } else {
        if (wide)
            hipLaunchKernelGGL(icephase_trace_drop_wide, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
        else
            hipLaunchKernelGGL(icephase_trace_drop, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    }
    return hipGetLastError();
}

// hex records traced from area-quadrature entries (DevRecord::area set)
hipError_t launch_trace_area(bool wide, const DevRecord* d_recs, TileMap map, uint32_t* d_queue, uint32_t blocks,
                             uint32_t threads, hipStream_t s)
{
    if (wide)
        hipLaunchKernelGGL(icephase_trace_hex_area_wide, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    else
        hipLaunchKernelGGL(icephase_trace_hex_area, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    return hipGetLastError();
}

hipError_t launch_bin_dirs(const BinFrame& g, uint32_t nt, uint32_t np, const double* d_edges, uint64_t n,
                           const double* d_dirs, uint32_t* d_index, int32_t* d_kind, hipStream_t s)
{
    uint64_t blocks = (n + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
    hipLaunchKernelGGL(icephase_bin_dirs, dim3((uint32_t)blocks), dim3(256), 0, s, g, nt, np, d_edges, n, d_dirs, d_index,
                       d_kind);
    return hipGetLastError();
}

}  // namespace icephase
