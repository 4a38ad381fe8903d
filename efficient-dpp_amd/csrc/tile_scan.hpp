// tile_scan.hpp — workgroup tile scan and the decoupled look-back protocol.
//
// Tile layout: a tile of TILE = BLOCK*EPV*VPT elements is VPT sub-tiles of
// BLOCK*EPV contiguous elements.  In sub-tile j lane t owns EPV consecutive
// elements, loaded with one 16-byte (dwordx4) access, so every load and store
// is coalesced and no LDS transpose is needed.  Prefix order inside the tile:
// sub-tile j, then wave w, then lane, then element.
#pragma once
#include "device.hpp"
#include "edpp.h"

namespace edpp {

enum : uint32_t { kFlagNone = 0, kFlagAggregate = 1, kFlagPrefix = 2 };

template <class T, int BLOCK, int VPT>
struct TileScan {
    typedef typename vec16<T>::type V;
    static constexpr int EPV = vec16<T>::n;
    static constexpr int SUB = BLOCK * EPV;
    static constexpr int TILE = SUB * VPT;
    static constexpr int NW = BLOCK / kWave;
    static_assert(BLOCK % kWave == 0, "block must be whole waves");

    struct Smem {
        T wave_tot[VPT][NW];  // wave totals, then their exclusive offsets in tile order
        T total;
        T prefix;
        uint32_t tile;
    };

    T v[VPT][EPV];
    T off[VPT];   // exclusive offset of the lane's first element of sub-tile j, within the tile
    T total;      // tile sum

    // The last, partial tile takes the same 16-byte vectors wherever a whole
    // vector lies below n; only the one vector that straddles n goes element
    // by element.  (A fully element-wise tail path holds 64-bit addresses for
    // every element at once and sizes the registers of the whole kernel.)
    __device__ __forceinline__ void load(const T* in, int64_t base, int64_t n) {
        const int t = threadIdx.x;
        if (base + TILE <= n) {
            #pragma unroll
            for (int j = 0; j < VPT; ++j) {
                V x = ld_stream(reinterpret_cast<const V*>(in + base + (int64_t)j * SUB) + t);
                #pragma unroll
                for (int k = 0; k < EPV; ++k) v[j][k] = x[k];
            }
        } else {
            #pragma unroll
            for (int j = 0; j < VPT; ++j) {
                const int64_t e0 = base + (int64_t)j * SUB + (int64_t)t * EPV;
                if (e0 + EPV <= n) {
                    V x = ld_stream(reinterpret_cast<const V*>(in + e0));
                    #pragma unroll
                    for (int k = 0; k < EPV; ++k) v[j][k] = x[k];
                } else {
                    #pragma unroll
                    for (int k = 0; k < EPV; ++k) v[j][k] = e0 + k < n ? in[e0 + k] : T(0);
                }
            }
        }
    }

    // Element-wise agent-scope atomic accesses: for data written by other
    // workgroups of the same launch, which may sit stale in another XCD's L2.
    __device__ __forceinline__ void load_coherent(const T* in, int64_t base, int64_t n) {
        const int t = threadIdx.x;
        #pragma unroll
        for (int j = 0; j < VPT; ++j)
            #pragma unroll
            for (int k = 0; k < EPV; ++k) {
                int64_t e = base + (int64_t)j * SUB + (int64_t)t * EPV + k;
                v[j][k] = e < n ? ld_relaxed(in + e) : T(0);
            }
    }
    __device__ __forceinline__ void store_coherent(T* out, int64_t base, int64_t n, T prefix) const {
        const int t = threadIdx.x;
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            T x = prefix + off[j];
            #pragma unroll
            for (int k = 0; k < EPV; ++k) {
                int64_t e = base + (int64_t)j * SUB + (int64_t)t * EPV + k;
                if (e < n) st_relaxed(out + e, x);
                x += v[j][k];
            }
        }
    }

    // Block-wide scan of the tile; contains two __syncthreads.  off[j] first
    // holds the lane's exclusive prefix inside its wave; wave 0 then turns the
    // VPT*NW wave totals into exclusive offsets in tile order (sub-tile, then
    // wave), and each lane adds the one offset of its wave to each off[j].
    // (Every lane summing all wave totals itself holds them all in registers.)
    __device__ __forceinline__ void scan(Smem& sm) {
        const int lane = lane_id();
        const int w = threadIdx.x / kWave;
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            T a = 0;
            #pragma unroll
            for (int k = 0; k < EPV; ++k) a += v[j][k];
            const T inc = wave_inclusive_add(a);
            off[j] = inc - a;
            if (lane == kWave - 1) sm.wave_tot[j][w] = inc;
        }
        __syncthreads();
        if (w == 0) {
            constexpr int M = VPT * NW;
            constexpr int PER = (M + kWave - 1) / kWave;  // consecutive entries per lane
            T* flat = &sm.wave_tot[0][0];
            T x[PER];
            T s = 0;
            #pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = lane * PER + i;
                x[i] = idx < M ? flat[idx] : T(0);
                s += x[i];
            }
            const T inc = wave_inclusive_add(s);
            T e = inc - s;
            #pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = lane * PER + i;
                if (idx < M) flat[idx] = e;
                e += x[i];
            }
            if (lane == kWave - 1) sm.total = inc;
        }
        __syncthreads();
        #pragma unroll
        for (int j = 0; j < VPT; ++j) off[j] += sm.wave_tot[j][w];
        total = sm.total;
    }

    // After scan(): turns v[j][k] into the element's exclusive prefix inside
    // the tile, so that off[] is dead and a tile needs VPT*EPV registers
    // through the look-back instead of VPT*(EPV+1); store_folded() then adds
    // the tile's prefix.
    __device__ __forceinline__ void fold() {
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            T x = off[j];
            #pragma unroll
            for (int k = 0; k < EPV; ++k) {
                const T a = v[j][k];
                v[j][k] = x;
                x += a;
            }
        }
    }
    __device__ __forceinline__ void store_folded(T* out, int64_t base, int64_t n, T prefix) const {
        const int t = threadIdx.x;
        const bool full = base + TILE <= n;
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            V o;
            #pragma unroll
            for (int k = 0; k < EPV; ++k) o[k] = prefix + v[j][k];
            const int64_t e0 = base + (int64_t)j * SUB + (int64_t)t * EPV;
            if (full || e0 + EPV <= n) {
                st_stream(o, reinterpret_cast<V*>(out + e0));
            } else {
                #pragma unroll
                for (int k = 0; k < EPV; ++k)
                    if (e0 + k < n) out[e0 + k] = o[k];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    __device__ __forceinline__ void store(T* out, int64_t base, int64_t n, T prefix) const {
        const int t = threadIdx.x;
        const bool full = base + TILE <= n;
        #pragma unroll
        for (int j = 0; j < VPT; ++j) {
            T x = prefix + off[j];
            V o;
            #pragma unroll
            for (int k = 0; k < EPV; ++k) { o[k] = x; x += v[j][k]; }
            const int64_t e0 = base + (int64_t)j * SUB + (int64_t)t * EPV;
            if (full || e0 + EPV <= n) {
                st_stream(o, reinterpret_cast<V*>(out + e0));
            } else {
                #pragma unroll
                for (int k = 0; k < EPV; ++k)
                    if (e0 + k < n) out[e0 + k] = o[k];
            }
            // one sub-tile's outputs at a time: its inputs die as its store issues
            __builtin_amdgcn_sched_barrier(0);
        }
    }
};

// Block-wide reduction helper (one __syncthreads); result valid in all lanes.
template <class T, int BLOCK>
__device__ __forceinline__ T block_sum(T x, T* lds /* BLOCK/64 entries */) {
    constexpr int NW = BLOCK / kWave;
    T ws = wave_sum(x);
    if (lane_id() == 0) lds[threadIdx.x / kWave] = ws;
    __syncthreads();
    T r = 0;
    #pragma unroll
    for (int q = 0; q < NW; ++q) r += lds[q];
    return r;
}

// ---------------------------------------------------------------- look-back state
// 32-bit values: one 64-bit word per tile = {epoch:30 | flag:2 | value:32}, so a
// single relaxed agent-scope load observes flag and value together.
struct LookbackState32 {
    uint64_t* status;
    uint32_t epoch;
    __device__ __forceinline__ void publish(int64_t t, uint32_t flag, uint32_t value) const {
        uint64_t w = ((uint64_t)epoch << 34) | ((uint64_t)flag << 32) | value;
        st_relaxed(status + t, w);
    }
    __device__ __forceinline__ void read(int64_t t, uint32_t& flag, uint32_t& value) const {
        uint64_t w = ld_relaxed(status + t);
        bool live = (uint32_t)(w >> 34) == epoch;
        flag = live ? (uint32_t)(w >> 32) & 3u : kFlagNone;
        value = (uint32_t)w;
    }
};

// 64-bit values: two 64-bit words per tile, each {epoch:30 | flag:2 | 32 bits
// of the value} (high half first).  Each word is single-copy atomic and
// carries its own tag, so a reader accepts the pair only when both words show
// this call's epoch and the same flag; a half-updated pair reads as "not yet"
// and is polled again.  No fence is needed on either side.
struct LookbackState64 {
    uint64_t* status;  // [2*t] high half, [2*t+1] low half
    uint32_t epoch;
    __device__ __forceinline__ void publish(int64_t t, uint32_t flag, uint64_t value) const {
        const uint64_t tag = ((uint64_t)epoch << 34) | ((uint64_t)flag << 32);
        st_relaxed(status + 2 * t, (uint64_t)(tag | (value >> 32)));
        st_relaxed(status + 2 * t + 1, (uint64_t)(tag | (value & 0xFFFFFFFFull)));
    }
    __device__ __forceinline__ void read(int64_t t, uint32_t& flag, uint64_t& value) const {
        const uint64_t hi = ld_relaxed(status + 2 * t);
        const uint64_t lo = ld_relaxed(status + 2 * t + 1);
        const bool live = (uint32_t)(hi >> 34) == epoch && (hi >> 32) == (lo >> 32);
        flag = live ? (uint32_t)(hi >> 32) & 3u : kFlagNone;
        value = (hi << 32) | (lo & 0xFFFFFFFFull);
    }
};

template <class T> struct LookbackFor;
template <> struct LookbackFor<uint32_t> { typedef LookbackState32 type; };
template <> struct LookbackFor<uint64_t> { typedef LookbackState64 type; };

// Decoupled look-back, executed by one whole wave for tile `tile` > 0.  One
// round trip reads kLookbackWindows windows of 64 predecessors (lane l of
// window j inspects tile cursor - 64*j - l); windows are consumed nearest
// first until a tile with an inclusive prefix is found.  Waiting only on
// smaller tickets, which belong to workgroups that are already running,
// guarantees progress.  Returns the exclusive prefix (valid in every lane).
// On watchdog expiry sets *err and returns what it has (the caller's result
// is then invalid and the host reports EDPP_ETIMEOUT).
constexpr int kLookbackWindows = 4;

template <class T, class State>
__device__ __forceinline__ T lookback(const State& st, int64_t tile, uint32_t spin_limit, int32_t* err) {
    const int lane = lane_id();
    T excl = 0;
    int64_t cursor = tile - 1;  // nearest predecessor not yet accounted for
    uint32_t spins = 0;
    while (true) {
        uint32_t flag[kLookbackWindows];
        T val[kLookbackWindows];
        #pragma unroll
        for (int j = 0; j < kLookbackWindows; ++j) {
            int64_t p = cursor - (int64_t)kWave * j - lane;
            if (p >= 0) {
                st.read(p, flag[j], val[j]);
            } else {
                flag[j] = kFlagPrefix;  // before tile 0: prefix 0
                val[j] = 0;
            }
        }
        bool stalled = false;
        #pragma unroll
        for (int j = 0; j < kLookbackWindows; ++j) {
            uint64_t pm = ballot(flag[j] == kFlagPrefix);
            uint64_t zm = ballot(flag[j] == kFlagNone);
            uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;  // lanes 0..first prefix lane
            if (zm & upto) {
                stalled = true;
                break;
            }
            excl += wave_sum(((upto >> lane) & 1ull) ? val[j] : T(0));
            if (pm) return excl;
            cursor -= kWave;
        }
        if (stalled) {
            if (++spins > spin_limit) {
                if (lane == 0) st_relaxed(err, (int32_t)EDPP_ETIMEOUT);
                return excl;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

}  // namespace edpp
