// fused_kernels.hpp -- two AA steps per launch on a periodic D3Q19 box (temporal blocking).
//
// The AA pattern's even step is local (a cell reads and writes its own slots) and its odd step
// reads and writes the neighbours' slots.  Fusing an even step with the following odd step lets
// the even step's post-collision populations stay on chip: a block owns a TX x TY tile of
// (x, y) columns plus a one-cell ring around it, marches up a chunk of z planes and, per plane,
//
//   1. runs the even step of plane p+1 on every cell of tile + ring (loads from `src`),
//   2. publishes to LDS what the odd step of plane p reads: the c_z = -1 populations of plane
//      p+1, the c_z = 0 populations of plane p and the c_z = +1 populations of plane p-1 (each
//      thread keeps its own column's later-needed populations in registers until then),
//   3. runs the odd step of plane p on the tile cells, gathering population v from the LDS copy
//      of cell x - c_v, and pushes its post-collision populations to `dst` at x + c_v.
//
// The ring costs a second even step on (TX+2)(TY+2)/(TX TY) - 1 of the cells and their reads.
// In return each cell is read once and written once per two steps instead of twice each, so
// the HBM traffic per two steps falls from 4 Q s to about (1 + ring) Q s + Q s bytes per cell.
// The odd step's pushes cross tile borders, so `dst` is a second population array: two
// launches in a row go src -> dst -> src.  Both arrays hold the state in the layout the even
// step reads (every cell's pre-collision populations in its own slots), which is the layout of
// a grid after a neighbour step (odd parity); a run from even parity takes one single
// neighbour step first.
//
// Every cell update is the same code as k_aa_periodic's (make_collider and the pair-wise
// collision), and the even step's populations go through the storage type S on the way through
// LDS, as they would through HBM, so in FP64 and FP32-shifted storage the result is bit for bit
// that of an even launch followed by an odd launch.  (In FP16 storage the compiler may round a
// collision's last FMA straight to half in one kernel and through float in the other: the two
// then differ by one unit in the last place of the stored half.)
#pragma once

#include "aa_kernels.hpp"

namespace lumagpu {

template <int TX, int TY>
struct FuseTile {
    static constexpr int WX = TX + 2, WY = TY + 2, C = WX * WY;
};

// Linear block id -> tile id such that the blocks one XCD receives (ids congruent mod 8 under the
// round-robin dispatch) take consecutive tiles: neighbouring tiles, whose ring rows overlap,
// then run on the same XCD and share its L2.
__device__ __forceinline__ int fused_tile_of(int b, int nblocks) {
    if (nblocks % 8) return b;
    return (b % 8) * (nblocks / 8) + b / 8;
}

// The 16 populations with a component in x or y, which the odd step takes from other columns
// (the other three, c = (0, 0, +-1) and the rest population, come from the thread's own column).
template <int V>
constexpr int lateral_index() {
    using L = Lattice<19>;
    int k = 0;
    for (int w = 0; w < V; ++w)
        if (L::c(w, 0) != 0 || L::c(w, 1) != 0) ++k;
    return (L::c(V, 0) != 0 || L::c(V, 1) != 0) ? k : -1;
}

// PF: the loads of plane p + 2 are issued at the start of plane p's iteration, into a second
// register set, so they arrive while plane p + 1's even step and plane p's odd step compute
// (one plane of load latency hidden per iteration instead of none).
// WIN: the LDS copy holds, per lateral population v, only the TX x TY window the tile reads
// (the value cell x + c_v of the tile pulls, stored at that cell's index): 16 * TX * TY values
// instead of 19 * (TX + 2)(TY + 2), so two copies of a 32 x 16 tile take 64 KiB and two blocks
// fit a CU; the odd step's LDS reads are then contiguous across the wave.
template <class P, int TX, int TY, bool STORE, bool DB, bool O32 = false, bool PF = false, bool WIN = false>
__global__ __launch_bounds__((TX + 2) * (TY + 2)) void k_aa_fused2(StepArgs a, void *dst_v, int zc) {
    using L = Lattice<19>;
    using S = typename P::S;
    using T = typename P::T;
    constexpr int Q = 19;
    constexpr int WX = FuseTile<TX, TY>::WX, C = FuseTile<TX, TY>::C;
    constexpr int TT = TX * TY;
    static_assert(!WIN || DB, "windowed LDS copies are double-buffered");
    // DB: two LDS copies used alternately, so a plane needs one barrier (between publishing and
    // reading) instead of two
    __shared__ S shm[WIN ? 2 * 16 * TT : (DB ? 2 : 1) * Q * C];

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int ntx = nx / TX, nty = ny / TY, ntz = (nz + zc - 1) / zc;
    const int tile = fused_tile_of(blockIdx.x, ntx * nty * ntz);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int t = threadIdx.x;
    // threads 0 .. TX*TY-1 take the tile row by row (whole waves, rows aligned to the tile's
    // first x), the rest the ring: the row below, the row above, then the two side columns
    int lx, ly;
    if (t < TX * TY) {
        lx = 1 + t % TX;
        ly = 1 + t / TX;
    } else if (t < TX * TY + WX) {
        lx = t - TX * TY;
        ly = 0;
    } else if (t < TX * TY + 2 * WX) {
        lx = t - TX * TY - WX;
        ly = TY + 1;
    } else {
        const int k = t - TX * TY - 2 * WX;
        lx = (k & 1) ? WX - 1 : 0;
        ly = 1 + (k >> 1);
    }
    const int lc = ly * WX + lx;   // the cell's slot in the LDS copy
    int gx = tx * TX - 1 + lx, gy = ty * TY - 1 + ly;
    gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
    gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
    const bool inner = t < TX * TY;
    const int z0 = tz * zc, z1 = min(z0 + zc, nz);

    const S *__restrict__ src = static_cast<const S *>(a.f);
    S *__restrict__ dst = static_cast<S *>(dst_v);
    const size_t ps = a.ps;
    const size_t col = (size_t)gx + (size_t)nx * (size_t)gy;
    const size_t plane = (size_t)nx * (size_t)ny;

    // the 19 populations of plane z of this thread's column, as stored
    auto load = [&](int z, S *raw) {
        const size_t cell = col + plane * (size_t)z;
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            if constexpr (O32)
                raw[v] = *reinterpret_cast<const S *>(reinterpret_cast<const char *>(src + (size_t)v * ps) +
                                                      (uint32_t)((uint32_t)cell * (uint32_t)sizeof(S)));
            else
                raw[v] = src[(size_t)v * ps + cell];
        });
    };
    // even step of a column cell from its loaded populations: post-collision populations in
    // storage precision
    auto even_from = [&](const S *raw, S *e) {
        T fin[Q];
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            fin[v] = static_cast<T>(raw[v]);
        });
        typename P::M rho, u[3];
        const auto c = make_collider<Q, P>(fin, nullptr, a, rho, u);
        static_for<(Q - 1) / 2>([&](auto pc) {
            constexpr int v = 2 * decltype(pc)::value;
            T ov, oo;
            c.template pair<v>(fin[v], fin[v + 1], ov, oo);
            e[v] = static_cast<S>(ov);
            e[v + 1] = static_cast<S>(oo);
        });
        e[Q - 1] = static_cast<S>(c(Q - 1, fin[Q - 1]));
    };
    auto even = [&](int z, S *e) {
        S raw[Q];
        load(z, raw);
        even_from(raw, e);
    };
    auto wrapz = [&](int z) { return z < 0 ? z + nz : (z >= nz ? z - nz : z); };

    // own column's populations awaiting publication: c_z = 0 of plane p, c_z = +1 of planes p
    // and p - 1 (indexed by population; only the entries of the matching group are used)
    S rmid[Q], rup[Q], rup_prev[Q];
    {
        S e[Q];
        even(wrapz(z0 - 1), e);
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            if constexpr (L::c(v, 2) == 1) rup_prev[v] = e[v];
        });
        even(z0, e);
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            if constexpr (L::c(v, 2) == 1) rup[v] = e[v];
            if constexpr (L::c(v, 2) == 0) rmid[v] = e[v];
        });
    }

    // one plane: the even step of plane p + 1 (its populations loaded into `cur`, or loaded
    // here), publication, the odd step of plane p.  With PF the loads of plane p + 2 go into
    // `nxt` first; the caller alternates two buffers, so no copy ties the loads to the loop end.
    auto plane_step = [&](int p, const S *cur, S *nxt) {
        S e[Q];
        if constexpr (PF) {
            // unconditionally (the last plane's load is spare): a branch here would make the
            // wait for `cur` also wait for these loads
            load(wrapz(p + 2), nxt);
            // keep the loads ahead of the plane's work (the scheduler would otherwise sink them
            // towards their use in the next plane)
            __builtin_amdgcn_sched_barrier(0);
            even_from(cur, e);
        } else {
            even(wrapz(p + 1), e);
        }
        S *sh = WIN ? shm + (p & 1) * (16 * TT) : (DB ? shm + (p & 1) * (Q * C) : shm);
        if constexpr (!DB) __syncthreads();   // the previous plane's odd step has read the LDS copy
        // the own-column populations of the odd step (WIN): c_z = +1 of plane p - 1, c_z = -1 of
        // plane p + 1, the rest population of plane p
        S own4 = S(0), own5 = S(0), own18 = S(0);
        if constexpr (WIN) {
            own4 = rup_prev[4];
            own5 = e[5];
            own18 = rmid[18];
        }
        static_for<Q>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            constexpr int cz = L::c(v, 2);
            S val;
            if constexpr (cz == -1) val = e[v];
            if constexpr (cz == 0) {
                val = rmid[v];
                rmid[v] = e[v];
            }
            if constexpr (cz == 1) {
                val = rup_prev[v];
                rup_prev[v] = rup[v];
                rup[v] = e[v];
            }
            if constexpr (!WIN) {
                sh[v * C + lc] = val;
            } else if constexpr (lateral_index<v>() >= 0) {
                // the tile cell that pulls this population from this cell
                const int ix = lx - 1 + L::c(v, 0), iy = ly - 1 + L::c(v, 1);
                if (ix >= 0 && ix < TX && iy >= 0 && iy < TY) sh[lateral_index<v>() * TT + iy * TX + ix] = val;
            }
        });
        __syncthreads();
        if (inner) {
            T fin[Q];
            static_for<Q>([&](auto vc) {
                constexpr int v = decltype(vc)::value;
                constexpr int cx = L::c(v, 0), cy = L::c(v, 1);
                if constexpr (!WIN) {
                    fin[v] = static_cast<T>(sh[v * C + lc - cy * WX - cx]);
                } else if constexpr (lateral_index<v>() >= 0) {
                    fin[v] = static_cast<T>(sh[lateral_index<v>() * TT + (ly - 1) * TX + (lx - 1)]);
                } else {
                    fin[v] = static_cast<T>(v == 4 ? own4 : (v == 5 ? own5 : own18));
                }
            });
            typename P::M rho, u[3];
            const auto c = make_collider<Q, P>(fin, nullptr, a, rho, u);
            if constexpr (STORE) store_macro<3>(a, col + plane * (size_t)p, rho, u);
            const int xm = gx == 0 ? nx - 1 : gx - 1, xp = gx == nx - 1 ? 0 : gx + 1;
            const int ym = gy == 0 ? ny - 1 : gy - 1, yp = gy == ny - 1 ? 0 : gy + 1;
            const int zm = p == 0 ? nz - 1 : p - 1, zp = p == nz - 1 ? 0 : p + 1;
            auto put = [&](auto vc, T o) {
                constexpr int v = decltype(vc)::value;
                constexpr int cx = L::c(v, 0), cy = L::c(v, 1), cz = L::c(v, 2);
                const int xx = cx == 0 ? gx : (cx > 0 ? xp : xm);
                const int yy = cy == 0 ? gy : (cy > 0 ? yp : ym);
                const int zz = cz == 0 ? p : (cz > 0 ? zp : zm);
                if constexpr (O32)
                    st<P, cx == 0>(elem_at<S, true>(dst + (size_t)v * ps,
                                                    (uint32_t)xx + (uint32_t)nx * ((uint32_t)yy + (uint32_t)ny * (uint32_t)zz)), o);
                else
                    st<P, cx == 0>(dst + (size_t)v * ps + cell_of(xx, yy, zz, nx, ny), o);
            };
            static_for<(Q - 1) / 2>([&](auto pc) {
                constexpr int v = 2 * decltype(pc)::value;
                T ov, oo;
                c.template pair<v>(fin[v], fin[v + 1], ov, oo);
                put(std::integral_constant<int, v>{}, ov);
                put(std::integral_constant<int, v + 1>{}, oo);
            });
            put(std::integral_constant<int, Q - 1>{}, c(Q - 1, fin[Q - 1]));
        }
    };
    if constexpr (PF) {
        S bufA[Q], bufB[Q];
        load(wrapz(z0 + 1), bufA);
        // pairs of planes with no branch inside the loop body (a conditional second plane merges
        // a path with bufB's loads still in flight into the loop header, whose wait would then
        // cover every outstanding load and store), then an odd last plane
        const int zp = z0 + ((z1 - z0) & ~1);
        for (int p = z0; p < zp; p += 2) {
            plane_step(p, bufA, bufB);
            plane_step(p + 1, bufB, bufA);
        }
        if (zp < z1) plane_step(zp, bufA, bufB);
    } else {
        for (int p = z0; p < z1; ++p) plane_step(p, nullptr, nullptr);
    }
}

// Tile shapes the launcher offers; the grid must be a whole number of tiles in x and y.  Shape
// codes: the tile (0-2), + 8 for two LDS copies, + 16 (with two copies) for loads one plane ahead.
enum FusedShape : int { kFuse64x8 = 0, kFuse32x16 = 1, kFuse128x4 = 2 };

inline bool fused_tile(int shape, int &tx, int &ty) {
    static const int t[3][2] = {{64, 8}, {32, 16}, {128, 4}};
    const int k = shape & 7;
    // tile (0-2), + 8 two LDS copies, + 16 loads one plane ahead, + 32 windowed copies; the
    // last two need the two copies
    if (shape < 0 || shape > 63 || k > 2 || (shape & 4) || ((shape & 48) && !(shape & 8))) return false;
    tx = t[k][0];
    ty = t[k][1];
    return true;
}

inline bool fused_shape_fits(int shape, int nx, int ny) {
    int tx, ty;
    if (!fused_tile(shape, tx, ty)) return false;
    return nx % tx == 0 && ny % ty == 0 && nx >= tx && ny >= ty;
}

template <class P, int TX, int TY, bool DB, bool PF = false, bool WIN = false>
hipError_t launch_fused2_t(const StepArgs &a, void *dst, int zc, bool store, hipStream_t s) {
    const int nblocks = (a.nx / TX) * (a.ny / TY) * ((a.nz + zc - 1) / zc);
    const dim3 g(nblocks), b(FuseTile<TX, TY>::C);
    // planes below 4 GiB: 32-bit byte offsets from each population plane's base
    // (measured at 256^3 FP32, 64x8: 302 us per step against 318 with 64-bit addresses)
    const bool o32 = a.ps * sizeof(typename P::S) < (size_t(1) << 32);
    if (o32) {
        if (store) k_aa_fused2<P, TX, TY, true, DB, true, PF, WIN><<<g, b, 0, s>>>(a, dst, zc);
        else k_aa_fused2<P, TX, TY, false, DB, true, PF, WIN><<<g, b, 0, s>>>(a, dst, zc);
        return hipGetLastError();
    }
    if (store) k_aa_fused2<P, TX, TY, true, DB, false, PF, WIN><<<g, b, 0, s>>>(a, dst, zc);
    else k_aa_fused2<P, TX, TY, false, DB, false, PF, WIN><<<g, b, 0, s>>>(a, dst, zc);
    return hipGetLastError();
}

template <class P>
hipError_t launch_fused2_p(int shape, const StepArgs &a, void *dst, int zc, bool store, hipStream_t s) {
    if (!fused_shape_fits(shape, a.nx, a.ny) || zc < 1 || a.nz < 3) return hipErrorInvalidValue;
    switch (shape) {
        case kFuse64x8: return launch_fused2_t<P, 64, 8, false>(a, dst, zc, store, s);
        case kFuse32x16: return launch_fused2_t<P, 32, 16, false>(a, dst, zc, store, s);
        case kFuse128x4: return launch_fused2_t<P, 128, 4, false>(a, dst, zc, store, s);
        default: break;
    }
    // two LDS copies fit the 160 KiB of a CU for storage types up to 4 bytes
    if constexpr (sizeof(typename P::S) <= 4) {
        if (shape == 8 + kFuse64x8) return launch_fused2_t<P, 64, 8, true>(a, dst, zc, store, s);
        if (shape == 8 + kFuse32x16) return launch_fused2_t<P, 32, 16, true>(a, dst, zc, store, s);
        if (shape == 8 + kFuse128x4) return launch_fused2_t<P, 128, 4, true>(a, dst, zc, store, s);
        if (shape == 24 + kFuse64x8) return launch_fused2_t<P, 64, 8, true, true>(a, dst, zc, store, s);
        if (shape == 24 + kFuse32x16) return launch_fused2_t<P, 32, 16, true, true>(a, dst, zc, store, s);
        if (shape == 24 + kFuse128x4) return launch_fused2_t<P, 128, 4, true, true>(a, dst, zc, store, s);
        if (shape == 40 + kFuse64x8) return launch_fused2_t<P, 64, 8, true, false, true>(a, dst, zc, store, s);
        if (shape == 40 + kFuse32x16) return launch_fused2_t<P, 32, 16, true, false, true>(a, dst, zc, store, s);
        if (shape == 56 + kFuse32x16) return launch_fused2_t<P, 32, 16, true, true, true>(a, dst, zc, store, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace lumagpu
