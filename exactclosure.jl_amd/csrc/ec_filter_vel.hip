// Velocity-type filters of several LES grids in one launch, for the RK4
// hot path of the DNS-aided LES:
//
//   out_s[a] = V_s(v_a - delta_a p)     (p optional)
//
// V_s is the filter that defines the LES velocity of set s: volume average
// (face top-hat / trapezoid along a, box over the other two axes) or face
// average (face point along a, box over the others), the pieces of
// apply_DHpi (src/N.jl:1228-1237) for any compression.  Applied to the
// projected DNS tendency k = du - grad p this gives V(k) = -div_H^ns (S r),
// the filtered flux divergence of get_rDHi (src/N.jl:1038-1073), so the exact
// closure forcing of the LES is available without forming the nine stresses.
//
// Structure (z-first separable filter): a block owns a TX x TX column tile
// and marches a chunk of ZC planes along z.  Each thread owns a few
// (component, x, y) columns and keeps, per set, the running z-piece of each
// column in an FP64 register, so a plane costs no LDS traffic and no
// barrier.  When the z-window of a coarse plane of some set closes (every c
// planes) the closed z-sums go to LDS and the block reduces them along x,
// then y, and writes the coarse values.  Component a needs columns up to c/2
// past the tile along a (its face piece) and none past it along the other
// axis, so no corner columns are read.  All sums are FP64.
#include "ec_internal.h"

#include <algorithm>
#include <numeric>
#include <type_traits>

namespace {

// periodic wrap of an index within one period of [0, n)
__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// 1D pieces over FP64 data: 0 box (cells cI..cI+c-1, i.e. f-c+1..f), 1 face
// (top-hat, or trapezoid for even c, around the coarse face f = c(I+1)-1),
// 2 point at f
template <int C, int PIECE>
__device__ __forceinline__ double psum(const double *p, int stride, int f) {
    if constexpr (PIECE == 0) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int t = 0; t < C; t++) {
            if (t & 1) a1 += p[(f - (C - 1) + t) * stride];
            else a0 += p[(f - (C - 1) + t) * stride];
        }
        return (a0 + a1) * (1.0 / C);
    } else if constexpr (PIECE == 1) {
        constexpr int H = C / 2;
        double a0 = 0.0, a1 = 0.0;
        if constexpr (C & 1) {
#pragma unroll
            for (int t = -H; t <= H; t++) {
                if (t & 1) a1 += p[(f + t) * stride];
                else a0 += p[(f + t) * stride];
            }
            return (a0 + a1) * (1.0 / C);
        } else {
#pragma unroll
            for (int t = -H + 1; t <= H - 1; t++) {
                if (t & 1) a1 += p[(f + t) * stride];
                else a0 += p[(f + t) * stride];
            }
            return ((a0 + a1) + 0.5 * (p[(f - H) * stride] + p[(f + H) * stride])) * (1.0 / C);
        }
    } else {
        return p[f * stride];
    }
}

// The split face piece of a volume-average set over the c cells of coarse
// cell J (p[0] its first fine cell, stride between cells): u = the part of
// face J's window (around the cell's last face) inside the cell, l = the part
// of face J-1's window (around the face below the cell) inside it, both / C.
// u + l is the box of the cell.  Odd C: windows of C faces; even C: C + 1
// faces with the two end faces at half weight, the middle face of the cell
// shared half and half.
template <int C>
__device__ __forceinline__ void half_sums(const double *p, int stride, double &u, double &l) {
    if constexpr (C & 1) {
        constexpr int H = (C - 1) / 2;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int t = 0; t < H; t++) b += p[t * stride];
#pragma unroll
        for (int t = H; t < C; t++) a += p[t * stride];
        u = a * (1.0 / C);
        l = b * (1.0 / C);
    } else {
        constexpr int M = C / 2 - 1;  // the shared middle face
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int t = 0; t < M; t++) b += p[t * stride];
#pragma unroll
        for (int t = M + 1; t < C; t++) a += p[t * stride];
        const double m = 0.5 * p[M * stride];
        u = (a + m) * (1.0 / C);
        l = (b + m) * (1.0 / C);
    }
}

struct FSet {
    double *out[3];
    double *side[2];  // volume average: the lower-half partials of components 0 and 1 (below)
    int nb;
};

struct FArgs {
    FSet s[4];
};

// Compile-time description of the tile (TX x TY columns, NT threads) and up
// to four sets (C = 0: absent) with their LDS layout.
template <int TX_, int TY_, int NT_, int C0, int K0, int C1, int K1, int C2, int K2, int C3, int K3>
struct Cfg {
    static constexpr int TX = TX_, TY = TY_, NT = NT_;
    static constexpr int c(int s) { return s == 0 ? C0 : s == 1 ? C1 : s == 2 ? C2 : C3; }
    static constexpr int kind(int s) { return s == 0 ? K0 : s == 1 ? K1 : s == 2 ? K2 : K3; }
    // planes needed past the chunk along z for component 2 (volume: face piece)
    static constexpr int hz(int s) { return c(s) && kind(s) == 0 ? c(s) / 2 : 0; }
    static constexpr int hmax = std::max(std::max(hz(0), hz(1)), std::max(hz(2), hz(3)));
    // columns past the tile along x / y: none.  The face pieces of components 0
    // and 1 of a volume-average set are split at the coarse-cell boundary
    // instead: each tile sums, per coarse cell J, the part of face J's window
    // inside cell J (U) and the part of face J-1's window inside cell J (L);
    // face J is U_J + L_{J+1}, added by a coarse-grid pass after the launch.
    static constexpr int h(int) { return 0; }
    static constexpr bool split(int s) { return c(s) && kind(s) == 0; }
    static constexpr int lcm() {
        int l = 1;
        for (int s = 0; s < 4; s++)
            if (c(s)) l = std::lcm(l, c(s));
        return l;
    }
    // columns of the tile: TX x TY per component; each region padded to whole
    // waves so a wave holds one component
    static constexpr int up64(int x) { return (x + 63) / 64 * 64; }
    static constexpr int N0 = TX * TY, N1 = TX * TY, N2 = TX * TY;
    static constexpr int P0 = up64(N0), P1 = up64(N1), P2 = up64(N2);
    static constexpr int NCC = P0 + P1 + P2;
    static constexpr int NJ = (NCC + NT - 1) / NT;
    // closed z-sums per (set, component): width (odd, for the banks) x height
    static constexpr int zw(int s, int a) { return (a == 0 ? TX + h(s) : TX) | 1; }
    static constexpr int zh(int s, int a) { return a == 1 ? TY + h(s) : TY; }
    static constexpr int zoff(int s, int a) {
        int o = 0;
        for (int t = 0; t < 4; t++)
            for (int b = 0; b < 3; b++) {
                if (t == s && b == a) return o;
                if (c(t)) o += zw(t, b) * zh(t, b);
            }
        return o;
    }
    // x-partials per (set, component): TX/c columns (two per coarse cell, U
    // and L, for the split component 0) of XR rows
    static constexpr int XR = TY | 1;
    static constexpr int xcols(int s, int a) { return (TX / c(s)) * (split(s) && a == 0 ? 2 : 1); }
    static constexpr int xoff(int s, int a) {
        int o = zoff(4, 0);
        for (int t = 0; t < 4; t++)
            for (int b = 0; b < 3; b++) {
                if (t == s && b == a) return o;
                if (c(t)) o += xcols(t, b) * XR;
            }
        return o;
    }
    static constexpr int lds_doubles = xoff(4, 0);
};

// z-window state of a set with compression C at plane r of the chunk
template <int C>
struct ZWin {
    bool box_on, box_start, box_emit;     // box / point windows [cK, cK + c - 1]
    bool face_on, face_start, face_emit;  // face windows around c(K + 1) - 1
    int kbox, kface;                      // chunk-relative coarse plane an emit closes
    __device__ __forceinline__ ZWin(int r, int ZC) {
        const int ph = r % C;
        box_on = r < ZC;
        box_start = ph == 0;
        box_emit = box_on && ph == C - 1;
        kbox = r / C;
        constexpr int r0 = (C & 1) ? (C - 1) / 2 : C / 2 - 1;
        const int d = r - r0;
        const int fph = d >= 0 ? d % C : 1;
        if constexpr (C & 1) {
            face_on = d >= 0 && d < ZC;
            face_start = fph == 0;
            face_emit = face_on && fph == C - 1;
            kface = d >= 0 ? d / C : 0;
        } else {
            // trapezoid: the last tap (weight 1/2) of window K is the first of K + 1
            face_on = d >= 0 && d <= ZC;
            face_start = fph == 0;
            face_emit = face_on && fph == 0 && d > 0;
            kface = d > 0 ? d / C - 1 : 0;
        }
    }
};

// z update of set S for one column of component a (uniform over the wave)
// with source value src.  The running sums are updated branch-free,
// acc = m * acc + w * src with uniform m (0 at a window start) and w: a
// window start discards what came before it, so planes outside every window
// of the chunk are harmless.  When a window closes, the closed z-sum goes to
// the LDS plane of (S, a); padding lanes (ok false) write nothing.
template <class G, int S>
__device__ __forceinline__ void zupdate(int r, int ZC, int a, bool ok, int lx, int ly, double src, double &acc,
                                        double *lds) {
    constexpr int C = G::c(S), K = G::kind(S), h = G::h(S);
    constexpr int r0 = (C & 1) ? (C - 1) / 2 : C / 2 - 1;  // first face-window start
    const int ph = r % C;
    const bool box_emit = r < ZC && ph == C - 1;
    if (a != 2 || K == 1) {
        if (a != 2) acc = (ph == 0 ? 0.0 : acc) + src;  // box along z
        if (box_emit && ok) {
            const double val = a == 2 ? src : acc * (1.0 / C);  // face average: point along z
            if (a == 0) {
                if (lx < G::TX + h) lds[G::zoff(S, 0) + ly * G::zw(S, 0) + lx] = val;
            } else if (a == 1) {
                if (ly < G::TY + h) lds[G::zoff(S, 1) + ly * G::zw(S, 1) + lx] = val;
            } else {
                lds[G::zoff(S, 2) + ly * G::zw(S, 2) + lx] = val;
            }
        }
    } else {  // volume average, component along z: face piece
        const int d = r - r0 + C;  // >= 1 for r >= 0; window starts where d % C == 0
        const int fph = d % C;
        if constexpr (C & 1) {
            acc = (fph == 0 ? 0.0 : acc) + src;
            if (fph == C - 1 && d - C < ZC && d >= C && ok)
                lds[G::zoff(S, 2) + ly * G::zw(S, 2) + lx] = acc * (1.0 / C);
        } else {
            // trapezoid: the last tap (weight 1/2) of window K is the first of K + 1
            if (fph == 0 && d > C && d - C <= ZC && ok)
                lds[G::zoff(S, 2) + ly * G::zw(S, 2) + lx] = (acc + 0.5 * src) * (1.0 / C);
            acc = fph == 0 ? 0.5 * src : acc + src;
        }
    }
}

// x stage of (S, A): x-partials of the closed z-sums, rows fastest
template <class G, int S, int A>
__device__ __forceinline__ void xstage(double *lds, int tid) {
    constexpr int C = G::c(S), K = G::kind(S);
    constexpr int rows = G::TY, nix = G::TX / C;
    constexpr int PX = A == 0 ? (K == 0 ? 1 : 2) : 0;
    const int t = (tid + 64 * (S * 3 + A)) % G::NT;  // spread the stages over the waves
    // at most one item per thread where the stage fits the block: a guarded
    // single iteration instead of a loop (see fv_march on wait counts)
    constexpr int STEP = nix * rows <= G::NT ? 0x7fffffff : G::NT;
    for (int it = t; it < nix * rows; it += STEP) {
        const int Ix = it / rows, row = it - Ix * rows;
        const double *zr = lds + G::zoff(S, A) + row * G::zw(S, A);
        if constexpr (A == 0 && G::split(S)) {
            double u, l;
            half_sums<C>(zr + C * Ix, 1, u, l);
            lds[G::xoff(S, A) + (2 * Ix) * G::XR + row] = u;
            lds[G::xoff(S, A) + (2 * Ix + 1) * G::XR + row] = l;
        } else {
            lds[G::xoff(S, A) + Ix * G::XR + row] = psum<C, PX>(zr, 1, C * (Ix + 1) - 1);
        }
    }
}

// y stage of (S, A): coarse values of the closed coarse plane Kc
template <class G, int S, int A>
__device__ __forceinline__ void ystage(const double *lds, int tid, int x0, int y0, int Kc, const FSet *fsl) {
    const FSet &fs = fsl[S];
    constexpr int C = G::c(S), K = G::kind(S);
    constexpr int nix = G::TX / C, niy = G::TY / C;
    constexpr int PY = A == 1 ? (K == 0 ? 1 : 2) : 0;
    const int t = (tid + 64 * (S * 3 + A)) % G::NT;
    constexpr int STEP = nix * niy <= G::NT ? 0x7fffffff : G::NT;
    for (int it = t; it < nix * niy; it += STEP) {
        const int Iy = it / nix, Ix = it - Iy * nix;
        const size_t nb = (size_t)fs.nb;
        const size_t J = (size_t)(x0 / C + Ix) + nb * ((size_t)(y0 / C + Iy) + nb * (size_t)Kc);
        if constexpr (A == 0 && G::split(S)) {
            // U and L of the x split, each boxed along y
            fs.out[0][J] = psum<C, 0>(lds + G::xoff(S, 0) + (2 * Ix) * G::XR, 1, C * (Iy + 1) - 1);
            fs.side[0][J] = psum<C, 0>(lds + G::xoff(S, 0) + (2 * Ix + 1) * G::XR, 1, C * (Iy + 1) - 1);
        } else if constexpr (A == 1 && G::split(S)) {
            // the face piece along y split at the coarse-cell boundary
            double u, l;
            half_sums<C>(lds + G::xoff(S, 1) + Ix * G::XR + C * Iy, 1, u, l);
            fs.out[1][J] = u;
            fs.side[1][J] = l;
        } else {
            fs.out[A][J] = psum<C, PY>(lds + G::xoff(S, A) + Ix * G::XR, 1, C * (Iy + 1) - 1);
        }
    }
}

template <class G, int S>
__device__ __forceinline__ void set_plane(int r, int ZC, const int *ca, const bool *ok, const int *clx,
                                          const int *cly, const double *src, double (*acc)[4], double *lds) {
    if constexpr (G::c(S) > 0) {
#pragma unroll
        for (int j = 0; j < G::NJ; j++) zupdate<G, S>(r, ZC, ca[j], ok[j], clx[j], cly[j], src[j], acc[j][S], lds);
    }
}

// The four-set list {16 volume, 8 volume, 4 volume, 4 face}: its box windows
// nest, so one running 4-plane box sum b4 per column serves every set.  At the
// last plane of 4-window m (phase 3): B = b4 closes the compression-4 box; the
// 8- and 16-plane boxes are sums of 2 and 4 consecutive B; the 8- and 16-plane
// trapezoids (ends of weight 1/2 at phase-3 planes) are
//   face8(K)  = x(2K)/2   + B(2K+1) + B(2K+2)                 - x(2K+2)/2
//   face16(K) = x(4K+1)/2 + B(4K+2) + ... + B(4K+5)            - x(4K+5)/2
// with x(m) the plane value at phase 3 of window m.  The compression-4
// trapezoid does not align with the 4-windows and keeps its own sum.
template <class G>
constexpr bool nested_list() {
    return G::c(0) == 16 && G::kind(0) == 0 && G::c(1) == 8 && G::kind(1) == 0 && G::c(2) == 4 &&
           G::kind(2) == 0 && G::c(3) == 4 && G::kind(3) == 1;
}

struct NestState {
    double b4, s1, s2, s3;  // components 0, 1: b4, box8, box16; component 2: b4, face4, face8, face16
    double hd;              // component 2: the last closed compression-4 trapezoid, written two planes later
};

template <class G, int S, int A>
__device__ __forceinline__ void zput(double *lds, int lx, int ly, double v) {
    lds[G::zoff(S, A) + ly * G::zw(S, A) + lx] = v;
}

// component A is needed up to h(S) columns past the tile along A (volume average)
template <class G, int S, int A>
__device__ __forceinline__ bool zin(int lx, int ly) {
    if constexpr (A == 0) return lx < G::TX + G::h(S);
    else if constexpr (A == 1) return ly < G::TY + G::h(S);
    else return true;
}

// components 0 and 1: box along z for every set (PH = r % 4, known at compile time)
template <class G, int A, int PH>
__device__ __forceinline__ void nested_box(int r, int ZC, bool ok, int lx, int ly, double src, NestState &st,
                                           double *lds) {
    const int m = r >> 2;
    st.b4 = PH == 0 ? src : st.b4 + src;
    if constexpr (PH == 3) {
        const double B = st.b4;
        st.s1 = ((m & 1) == 0 ? 0.0 : st.s1) + B;
        st.s2 = ((m & 3) == 0 ? 0.0 : st.s2) + B;
        if (r < ZC && ok) {
            if (zin<G, 3, A>(lx, ly)) zput<G, 3, A>(lds, lx, ly, B * 0.25);
            if (zin<G, 2, A>(lx, ly)) zput<G, 2, A>(lds, lx, ly, B * 0.25);
            if ((m & 1) == 1 && zin<G, 1, A>(lx, ly)) zput<G, 1, A>(lds, lx, ly, st.s1 * 0.125);
            if ((m & 3) == 3 && zin<G, 0, A>(lx, ly)) zput<G, 0, A>(lds, lx, ly, st.s2 * 0.0625);
        }
    }
}

// component 2: face average point, compression-4 trapezoid on its own sum,
// compression-8 and -16 trapezoids from the 4-plane boxes
template <class G, int PH>
__device__ __forceinline__ void nested_face(int r, int ZC, bool ok, int lx, int ly, double src, NestState &st,
                                            double *lds) {
    const int m = r >> 2;
    st.b4 = PH == 0 ? src : st.b4 + src;
    // compression-4 trapezoid: windows start at planes 4K + 1 (phase 1); the
    // closed sum waits in hd for the next phase-3 plane, so that every set
    // emits on the phase-3 planes only (one barrier pair per 4 planes)
    if constexpr (PH == 1) {
        st.hd = (st.s1 + 0.5 * src) * 0.25;
        st.s1 = 0.5 * src;
    } else {
        st.s1 += src;
    }
    if constexpr (PH == 3) {
        const double B = st.b4, hx = 0.5 * src;
        if (r > 3 && r - 3 <= ZC && ok) zput<G, 2, 2>(lds, lx, ly, st.hd);  // closed at plane r - 2
        if (r < ZC && ok) zput<G, 3, 2>(lds, lx, ly, src);
        if ((m & 1) == 0) {
            if (m > 0 && m <= ZC / 4 && ok) zput<G, 1, 2>(lds, lx, ly, (st.s2 + B - hx) * 0.125);
            st.s2 = hx;
        } else {
            st.s2 += B;
        }
        if ((m & 3) == 1) {
            if (m >= 5 && m <= ZC / 4 + 1 && ok) zput<G, 0, 2>(lds, lx, ly, (st.s3 + B - hx) * 0.0625);
            st.s3 = hx;
        } else {
            st.s3 += B;
        }
    }
}

// plane whose face window is emitted at plane r: the compression-4 trapezoid of
// the nested list is written two planes after it closes (see nested_face)
template <class G, int S>
__device__ __forceinline__ int face_plane(int r) {
    return nested_list<G>() && S == 2 ? r - 2 : r;
}

template <class G, int S>
__device__ __forceinline__ bool set_emits(int r, int ZC) {
    if constexpr (G::c(S) > 0) {
        const ZWin<G::c(S)> w(r, ZC), wf(face_plane<G, S>(r), ZC);
        return w.box_emit || (G::kind(S) == 0 && wf.face_emit);
    } else {
        return false;
    }
}

template <class G, int S>
__device__ __forceinline__ void set_xstage(int r, int ZC, double *lds, int tid) {
    if constexpr (G::c(S) > 0) {
        const ZWin<G::c(S)> w(r, ZC), wf(face_plane<G, S>(r), ZC);
        if (w.box_emit) {
            xstage<G, S, 0>(lds, tid);
            xstage<G, S, 1>(lds, tid);
            if constexpr (G::kind(S) == 1) xstage<G, S, 2>(lds, tid);
        }
        if constexpr (G::kind(S) == 0)
            if (wf.face_emit) xstage<G, S, 2>(lds, tid);
    }
}

template <class G, int S>
__device__ __forceinline__ void set_ystage(int r, int ZC, const double *lds, int tid, int x0, int y0, int z0,
                                           const FSet *fs) {
    if constexpr (G::c(S) > 0) {
        constexpr int C = G::c(S);
        const ZWin<C> w(r, ZC), wf(face_plane<G, S>(r), ZC);
        if (w.box_emit) {
            ystage<G, S, 0>(lds, tid, x0, y0, z0 / C + w.kbox, fs);
            ystage<G, S, 1>(lds, tid, x0, y0, z0 / C + w.kbox, fs);
            if constexpr (G::kind(S) == 1) ystage<G, S, 2>(lds, tid, x0, y0, z0 / C + w.kbox, fs);
        }
        if constexpr (G::kind(S) == 0)
            if (wf.face_emit) ystage<G, S, 2>(lds, tid, x0, y0, z0 / C + wf.kface, fs);
    }
}

// Component of column slot j of wave w (3: a padding slot).  The component
// regions are whole waves, so the component is uniform over a wave; every
// wave's slot components form a code of 2 bits per slot.
template <class G>
constexpr int slot_comp(int w, int j) {
    const int e = w * 64 + j * G::NT;
    return e < G::P0 ? 0 : (e < G::P0 + G::P1 ? 1 : (e < G::NCC ? 2 : 3));
}
template <class G>
constexpr int wave_code(int w) {
    int c = 0;
    for (int j = 0; j < G::NJ; j++) c |= slot_comp<G>(w, j) << (2 * j);
    return c;
}
template <class G>
constexpr bool first_code(int w) {
    for (int v = 0; v < w; v++)
        if (wave_code<G>(v) == wave_code<G>(w)) return false;
    return true;
}

struct FvTile {
    int n, ZC, x0, y0, z0;
    double ih;
};

// The DNS RK4 stage fused into the march (ec_ns_filter_vel_rk4): the march
// already forms k_a = du_a - delta_a p for every column, so the interior cells
// of the tile also take the stage update of ec_ns_rk4_update (same arithmetic,
// same FP32 k), and du and p are read once per stage instead of twice.
// RK = 0..3: stage RK with the accumulator acc (ec_ns_rk4_update);
// RK = 4..7: stage RK - 4 of the stored-state form (ec_ns_rk4_update_states),
// whose stages 0..2 write u1, u2, u3 (ust) and whose last stage combines them.
struct RkArgs {
    void *u0[3], *us[3], *acc[3];
    void *ust[3][3];
    double dt;
};

// values of the stage state read per cell: u0, with acc after stage 0 of the
// accumulator form, with u1, u2, u3 at the last stage of the stored-state form
template <int RK>
constexpr int rk_nv() {
    return RK == 7 ? 4 : ((RK >= 1 && RK <= 3) ? 2 : 1);
}

// loads of the stage state of one cell, issued at the top of a plane so that
// they arrive during the plane's filter work
// (read once: non-temporal, so that L2 keeps the pressure lines the
// neighbouring tiles share)
template <typename T, int RK>
__device__ __forceinline__ void rk_load(int a, unsigned I, const RkArgs &rk, T *v) {
    v[0] = __builtin_nontemporal_load((const T *)rk.u0[a] + I);
    if constexpr (RK >= 1 && RK <= 3) v[1] = __builtin_nontemporal_load((const T *)rk.acc[a] + I);
    if constexpr (RK == 7) {
        v[1] = __builtin_nontemporal_load((const T *)rk.ust[0][a] + I);
        v[2] = __builtin_nontemporal_load((const T *)rk.ust[1][a] + I);
        v[3] = __builtin_nontemporal_load((const T *)rk.ust[2][a] + I);
    }
}
template <typename T, int RK>
__device__ __forceinline__ void rk_store(int a, unsigned I, T k, const T *v, const RkArgs &rk) {
    const T dt = (T)rk.dt;
    const T u = v[0];
    if constexpr (RK == 0) {
        ec_nt_store((T *)rk.acc[a] + I, k);
        ec_nt_store((T *)rk.us[a] + I, u + (dt / 2) * k);
    } else if constexpr (RK == 1 || RK == 2) {
        ec_nt_store((T *)rk.acc[a] + I, v[1] + 2 * k);
        ec_nt_store((T *)rk.us[a] + I, u + (RK == 1 ? dt / 2 : dt) * k);
    } else if constexpr (RK == 3) {
        ec_nt_store((T *)rk.u0[a] + I, u + (dt / 6) * (v[1] + k));
    } else if constexpr (RK < 7) {
        ec_nt_store((T *)rk.ust[RK - 4][a] + I, u + ec_rk4_stage_coef<T>(RK - 4, dt) * k);
    } else {
        ec_nt_store((T *)rk.u0[a] + I, ec_rk4_combine<T>(u, v[1], v[2], v[3], k, dt));
    }
}

// The plane march of one wave whose slot components are CODE: component,
// column and window logic are compile-time, so a plane is straight-line code
// (loads, FP64 z-updates) up to the uniform emit branch.
// Element offsets are 32-bit: the host requires n^3 < 2^32.
// The plane body has no branch around a global load: the pressure's presence
// is a template flag, slots are known valid where the tile has no padding,
// and stage-state loads past the chunk read a wrapped plane (unused).  With a
// load under a branch, the compiler's wait-count insertion met a join it could
// not resolve and waited for every outstanding load (vmcnt(0)) before the
// plane's arithmetic, the next plane's prefetch included.
template <typename T, class G, int CODE, int RK, bool HASP>
__device__ __forceinline__ void fv_march(const T *__restrict__ v1, const T *__restrict__ v2, const T *__restrict__ v3,
                                         const T *__restrict__ p, const FvTile &tl, double *lds, const FSet *fsl,
                                         int tid, const RkArgs &rk) {
    constexpr int TX = G::TX, TY = G::TY, H = G::hmax, NJ = G::NJ, NT = G::NT;
    const int n = tl.n, ZC = tl.ZC, x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
    const unsigned nn = (unsigned)n, n2 = nn * nn;
    auto comp = [](int j) constexpr { return (CODE >> (2 * j)) & 3; };

    // the thread's columns: tile-local position, offsets of the value and of
    // the pressure neighbour along the component
    int clx[NJ], cly[NJ];
    bool ok[NJ];
    unsigned co[NJ], cq[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int a = comp(j);
        int e = tid + NT * j, lx = 0, ly = 0;
        bool valid = false;
        if (a == 0) {
            valid = e < G::N0, lx = e % TX, ly = e / TX;
        } else if (a == 1) {
            e -= G::P0;
            valid = e < G::N1, lx = e % TX, ly = e / TX;
        } else if (a == 2) {
            e -= G::P0 + G::P1;
            valid = e < G::N2, lx = e % TX, ly = e / TX;
        }
        if (!valid) lx = ly = 0;  // padding: reads column (0, 0), writes nothing
        ok[j] = G::NCC == G::NJ * G::NT && G::P0 == G::N0 && G::P1 == G::N1 && G::P2 == G::N2 ? true : valid;
        clx[j] = lx;
        cly[j] = ly;
        const unsigned xo = (unsigned)wrap1(x0 + lx, n), yo = (unsigned)wrap1(y0 + ly, n);
        co[j] = xo + nn * yo;
        cq[j] = a == 0 ? (unsigned)wrap1(x0 + lx + 1, n) + nn * yo
                       : (a == 1 ? xo + nn * (unsigned)wrap1(y0 + ly + 1, n) : co[j]);
    }

    constexpr bool NEST = nested_list<G>();
    double acc[NEST ? 1 : NJ][4];
    NestState nst[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        nst[j] = NestState{0.0, 0.0, 0.0, 0.0, 0.0};
        if constexpr (!NEST)
#pragma unroll
            for (int s = 0; s < 4; s++) acc[j][s] = 0.0;
    }

    // register double buffer: v, p and the p neighbour of plane r
    T cv[NJ], cp[NJ], cqv[NJ];
    // HALO: a plane past the chunk, which only component 2's face windows read
    // (the box windows of components 0 and 1 close inside the chunk)
    auto load = [&](int r, T *vv, T *pp, T *qq, auto halo) {
        constexpr bool HALO = decltype(halo)::value;
        const unsigned kk = (unsigned)wrap1(z0 + r, n) * n2, k1 = (unsigned)wrap1(z0 + r + 1, n) * n2;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int a = comp(j);
            if (a == 3) continue;
            if (HALO && a != 2) {  // (defined values, never read)
                vv[j] = pp[j] = qq[j] = T(0);
                continue;
            }
            const T *vp = a == 0 ? v1 : (a == 1 ? v2 : v3);
            // the filtered field is read once (non-temporal); the pressure and
            // its x/y neighbours are shared with the neighbouring tiles
            vv[j] = __builtin_nontemporal_load(vp + co[j] + kk);
            if constexpr (HASP) {
                pp[j] = p[co[j] + kk];
                qq[j] = p[cq[j] + (a == 2 ? k1 : kk)];
            } else {
                pp[j] = qq[j] = T(0);
            }
        }
    };
    load(0, cv, cp, cqv, std::false_type{});
    // fused RK4 stage: the stage state of plane r (rk_nv values per cell) is
    // loaded one plane ahead, with the filter's loads of that plane
    constexpr int NV = rk_nv<RK>();
    T rkv[NJ][NV];
    auto rk_prefetch = [&](int r, T (*vv)[NV]) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
            if (comp(j) != 3) rk_load<T, RK>(comp(j), co[j] + (unsigned)wrap1(z0 + r, n) * n2, rk, vv[j]);
    };
    auto rk_rotate = [&](T (*vv)[NV]) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int q = 0; q < NV; q++) rkv[j][q] = vv[j][q];
    };
    if constexpr (RK >= 0) rk_prefetch(0, rkv);
    const int rend = ZC - 1 + H;
    // k = v - delta p in the field precision (one rounding at its scale); the
    // z-sums and everything after them are FP64
    const T ihT = (T)tl.ih;
    // emits of every set at plane r (all on phase-3 planes for the nested list)
    auto emit = [&](int r) {
        const bool any = set_emits<G, 0>(r, ZC) || set_emits<G, 1>(r, ZC) || set_emits<G, 2>(r, ZC) ||
                         set_emits<G, 3>(r, ZC);
        if (any) {
            // an opaque copy of the thread index keeps the stage indexing from
            // being hoisted out of the plane loop (it would hold registers)
            int t = tid;
            asm volatile("" : "+v"(t));
            __syncthreads();
            set_xstage<G, 0>(r, ZC, lds, t);
            set_xstage<G, 1>(r, ZC, lds, t);
            set_xstage<G, 2>(r, ZC, lds, t);
            set_xstage<G, 3>(r, ZC, lds, t);
            __syncthreads();
            set_ystage<G, 0>(r, ZC, lds, t, x0, y0, z0, fsl);
            set_ystage<G, 1>(r, ZC, lds, t, x0, y0, z0, fsl);
            set_ystage<G, 2>(r, ZC, lds, t, x0, y0, z0, fsl);
            set_ystage<G, 3>(r, ZC, lds, t, x0, y0, z0, fsl);
        }
    };
    auto rotate = [&](const T *nv, const T *np, const T *nq) {
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            cv[j] = nv[j];
            cp[j] = np[j];
            cqv[j] = nq[j];
        }
    };
    if constexpr (NEST) {
        // four planes per iteration (ZC + H is a multiple of 4): the window
        // phase of each is a compile-time constant, and only the phase-3 plane
        // can emit
        static_assert(G::hmax % 4 == 0 && G::lcm() % 4 == 0, "nested march: whole 4-plane groups");
        // the chunk's planes (ZC, a multiple of 4) take the stage update; the
        // halo planes after them (to ZC + hmax) only feed the face windows, so
        // the plane body needs no per-plane store condition
        auto group = [&](int r0, auto st) {
            constexpr bool STORE = decltype(st)::value;
            auto plane = [&](auto qc) {
                constexpr int Q = decltype(qc)::value;
                const int r = r0 + Q;
                // plane r + 1 (its filter loads and stage state) in flight through this plane
                T nv[NJ], np[NJ], nq[NJ];
                load(r + 1, nv, np, nq, std::integral_constant<bool, !STORE>{});
                T rkn[NJ][NV];
                if constexpr (RK >= 0 && STORE) rk_prefetch(r + 1, rkn);
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int a = comp(j);
                    if (a == 3 || (!STORE && a != 2)) continue;
                    const T kv = cv[j] - (cqv[j] - cp[j]) * ihT;
                    const double src = (double)kv;
                    if constexpr (RK >= 0 && STORE)
                        if (ok[j]) rk_store<T, RK>(a, co[j] + (unsigned)(z0 + r) * n2, kv, rkv[j], rk);
                    if (a == 0) nested_box<G, 0, Q>(r, ZC, ok[j], clx[j], cly[j], src, nst[j], lds);
                    else if (a == 1) nested_box<G, 1, Q>(r, ZC, ok[j], clx[j], cly[j], src, nst[j], lds);
                    else nested_face<G, Q>(r, ZC, ok[j], clx[j], cly[j], src, nst[j], lds);
                }
                if constexpr (Q == 3) emit(r);
                rotate(nv, np, nq);
                if constexpr (RK >= 0 && STORE) rk_rotate(rkn);
            };
            plane(std::integral_constant<int, 0>{});
            plane(std::integral_constant<int, 1>{});
            plane(std::integral_constant<int, 2>{});
            plane(std::integral_constant<int, 3>{});
        };
        for (int r0 = 0; r0 < ZC; r0 += 4) group(r0, std::true_type{});
        for (int r0 = ZC; r0 <= rend; r0 += 4) group(r0, std::false_type{});
    } else {
        for (int r = 0; r <= rend; ++r) {
            T nv[NJ], np[NJ], nq[NJ];
            load(r + 1, nv, np, nq, std::false_type{});
            T rkn[NJ][NV];
            if constexpr (RK >= 0) rk_prefetch(r + 1, rkn);
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const int a = comp(j);
                if (a == 3) continue;
                const T kv = cv[j] - (cqv[j] - cp[j]) * ihT;
                const double src = (double)kv;
                if constexpr (RK >= 0)
                    if (r < ZC && ok[j])
                        rk_store<T, RK>(a, co[j] + (unsigned)(z0 + r) * n2, kv, rkv[j], rk);
                if constexpr (G::c(0) > 0) zupdate<G, 0>(r, ZC, a, ok[j], clx[j], cly[j], src, acc[j][0], lds);
                if constexpr (G::c(1) > 0) zupdate<G, 1>(r, ZC, a, ok[j], clx[j], cly[j], src, acc[j][1], lds);
                if constexpr (G::c(2) > 0) zupdate<G, 2>(r, ZC, a, ok[j], clx[j], cly[j], src, acc[j][2], lds);
                if constexpr (G::c(3) > 0) zupdate<G, 3>(r, ZC, a, ok[j], clx[j], cly[j], src, acc[j][3], lds);
            }
            emit(r);
            rotate(nv, np, nq);
            if constexpr (RK >= 0) rk_rotate(rkn);
        }
    }
}

// one instantiation of the march per distinct slot-component code
template <typename T, class G, int RK, bool HASP, int W>
__device__ __forceinline__ void fv_dispatch(int code, const T *__restrict__ v1, const T *__restrict__ v2,
                                            const T *__restrict__ v3, const T *__restrict__ p, const FvTile &tl,
                                            double *lds, const FSet *fsl, int tid, const RkArgs &rk) {
    if constexpr (W < G::NT / 64) {
        if constexpr (first_code<G>(W)) {
            if (code == wave_code<G>(W)) {
                fv_march<T, G, wave_code<G>(W), RK, HASP>(v1, v2, v3, p, tl, lds, fsl, tid, rk);
                return;
            }
        }
        fv_dispatch<T, G, RK, HASP, W + 1>(code, v1, v2, v3, p, tl, lds, fsl, tid, rk);
    }
}

// RK < 0: the filter alone; RK = 0..7: with the fused DNS RK4 stage (RkArgs).
// HASP: the source is v - delta p (else v).
template <typename T, class G, int RK, bool HASP = true>
__global__ void __launch_bounds__(G::NT) k_filter_vel(const T *__restrict__ v1, const T *__restrict__ v2,
                                                     const T *__restrict__ v3, const T *__restrict__ p, int n, int ZC,
                                                     double ih, FArgs A, RkArgs R) {
    constexpr int TX = G::TX, TY = G::TY;
    extern __shared__ double lds[];
    // output table in LDS: read only when a coarse plane closes, so it holds
    // no scalar registers through the plane loop
    FSet *fsl = reinterpret_cast<FSet *>(lds + G::lds_doubles);
    if (threadIdx.x < 4) fsl[threadIdx.x] = A.s[threadIdx.x];
    // 1D grid of (n/TX) x (n/TY) x (n/ZC) tiles, x fastest, in XCD order
    const int gx = n / TX, gy = n / TY;
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    FvTile tl;
    tl.n = n;
    tl.ZC = ZC;
    tl.ih = ih;
    tl.x0 = (L % gx) * TX;
    tl.y0 = ((L / gx) % gy) * TY;
    tl.z0 = (L / (gx * gy)) * ZC;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int code = 0;
#pragma unroll
    for (int j = 0; j < G::NJ; j++) {
        const int e = w * 64 + j * G::NT;
        const int c = e < G::P0 ? 0 : (e < G::P0 + G::P1 ? 1 : (e < G::NCC ? 2 : 3));
        code |= c << (2 * j);
    }
    fv_dispatch<T, G, RK, HASP, 0>(code, v1, v2, v3, p, tl, lds, fsl, tid, R);
}

// Face J of a split component = its U partial (already in out) + the L
// partial of the next coarse cell along the component's axis (periodic).
__global__ void k_face_combine(int nb, double *__restrict__ out0, const double *__restrict__ side0,
                               double *__restrict__ out1, const double *__restrict__ side1) {
    const size_t n1 = (size_t)nb, total = n1 * n1 * n1;
    const size_t J = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (J >= total) return;
    const int i = (int)(J % n1), j = (int)((J / n1) % n1);
    const size_t k = J / (n1 * n1);
    const int ip = i + 1 == nb ? 0 : i + 1, jp = j + 1 == nb ? 0 : j + 1;
    out0[J] += side0[(size_t)ip + n1 * ((size_t)j + n1 * k)];
    out1[J] += side1[(size_t)i + n1 * ((size_t)jp + n1 * k)];
}

int largest_tile(int n, int L, int target) {
    int best = 0;
    for (int t = L; t <= n; t += L)
        if (n % t == 0 && t <= target) best = t;
    if (best == 0)
        for (int t = L; t <= n; t += L)
            if (n % t == 0) return t;
    return best;
}

template <typename T, class G, int RK = -1>
int launch_fv(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, const FArgs &A, int tz_target,
              const RkArgs *rk = nullptr) {
    constexpr int TX = G::TX, TY = G::TY, L = G::lcm();
    static_assert(TX % L == 0 && TY % L == 0, "tile edges must be multiples of every compression");
    if (n % TX || n % TY || n % L) return EC_E_UNSUPPORTED;
    if ((double)n * n * n >= 4294967296.0) return EC_E_UNSUPPORTED;  // 32-bit element offsets
    // every index read (planes up to z0 + ZC + hmax + 1, columns up to x0 + TX,
    // y0 + TY) lies within one period
    if (G::hmax + 2 >= n) return EC_E_UNSUPPORTED;
    const int ZC = largest_tile(n, L, tz_target);
    if (!ZC) return EC_E_UNSUPPORTED;
    const size_t lds = sizeof(double) * (size_t)G::lds_doubles + sizeof(FSet) * 4;
    if (lds > 150 * 1024) return EC_E_UNSUPPORTED;
    const long tiles = (long)(n / TX) * (n / TY) * (n / ZC);
    if (tiles > 0x7fffffffL) return EC_E_UNSUPPORTED;
    // fused stage: reads u0 (and acc after stage 0), writes us and acc (u0 at
    // stage 3); stored-state form: reads u0, writes one state, or reads u0, u1,
    // u2, u3 and writes u0 at the last stage
    constexpr double rk_units =
        RK < 0 ? 0.0 : (RK == 0 || RK == 3 ? 9.0 : (RK <= 2 ? 12.0 : (RK == 7 ? 15.0 : 6.0)));
    ec_prof_scope pr(ctx, RK < 0 ? "filter_vel" : "filter_vel_rk4",
                     sizeof(T) * ((p ? 4.0 : 3.0) + rk_units) * ec_cube(n));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_filter_vel<T, G, RK, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_filter_vel<T, G, RK, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    RkArgs R{};
    if (rk) R = *rk;
    // the L partials of the split volume-average sets, in a context workspace
    FArgs AA = A;
    size_t need = 0;
    for (int s = 0; s < 4; s++)
        if (G::split(s)) need += 2 * ec_cube(AA.s[s].nb);
    if (need > ctx->ws3_doubles) {
        if (ctx->ws3) {
            EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            EC_HIP(ctx, hipFree(ctx->ws3));
            ctx->ws3 = nullptr;
            ctx->ws3_doubles = 0;
        }
        if (hipMalloc(&ctx->ws3, sizeof(double) * need) != hipSuccess)
            return ec_fail(ctx, EC_E_OOM, "velocity filter: %zu doubles of partials", need);
        ctx->ws3_doubles = need;
    }
    double *side = (double *)ctx->ws3;
    for (int s = 0; s < 4; s++) {
        AA.s[s].side[0] = AA.s[s].side[1] = nullptr;
        if (G::split(s)) {
            const size_t m = ec_cube(AA.s[s].nb);
            AA.s[s].side[0] = side;
            AA.s[s].side[1] = side + m;
            side += 2 * m;
        }
    }
    if (p)
        k_filter_vel<T, G, RK, true><<<dim3((unsigned)tiles), G::NT, lds, ctx->stream>>>(
            (const T *)v[0], (const T *)v[1], (const T *)v[2], (const T *)p, n, ZC, n / l, AA, R);
    else
        k_filter_vel<T, G, RK, false><<<dim3((unsigned)tiles), G::NT, lds, ctx->stream>>>(
            (const T *)v[0], (const T *)v[1], (const T *)v[2], nullptr, n, ZC, n / l, AA, R);
    EC_CHECK_LAUNCH(ctx, "k_filter_vel");
    for (int s = 0; s < 4; s++) {
        if (!G::split(s)) continue;
        const size_t m = ec_cube(AA.s[s].nb);
        k_face_combine<<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(AA.s[s].nb, AA.s[s].out[0], AA.s[s].side[0],
                                                                              AA.s[s].out[1], AA.s[s].side[1]);
        EC_CHECK_LAUNCH(ctx, "k_face_combine");
    }
    return EC_OK;
}

// one set with compression c (instantiated for the common compressions)
template <typename T, int K>
int launch_single(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, const FArgs &A, int c, int tz) {
    switch (c) {
        case 2: return launch_fv<T, Cfg<32, 16, 512, 2, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        case 3: return launch_fv<T, Cfg<24, 12, 256, 3, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        case 4: return launch_fv<T, Cfg<32, 16, 512, 4, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        case 5: return launch_fv<T, Cfg<20, 10, 256, 5, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        case 8: return launch_fv<T, Cfg<32, 16, 512, 8, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        case 16: return launch_fv<T, Cfg<32, 16, 512, 16, K, 0, 0, 0, 0, 0, 0>>(ctx, n, l, v, p, A, tz);
        default: return EC_E_UNSUPPORTED;
    }
}

// The four-set list with the DNS RK4 stage fused (one launch), or
// EC_E_UNSUPPORTED when the list or the grid takes another path.
template <typename T>
int filter_vel_rk4_t(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, int nsets,
                     const ec_vel_set *sets, int tz, int stage, const RkArgs &rk, bool states) {
    if (nsets != 4 || !p) return EC_E_UNSUPPORTED;
    FArgs A;
    int cs[4], ks[4];
    for (int s = 0; s < 4; s++) {
        cs[s] = n / sets[s].nbar;
        ks[s] = sets[s].kind;
        for (int a = 0; a < 3; a++) A.s[s].out[a] = (double *)sets[s].out[a];
        A.s[s].nb = sets[s].nbar;
    }
    if (!(cs[0] == 16 && ks[0] == 0 && cs[1] == 8 && ks[1] == 0 && cs[2] == 4 && ks[2] == 0 && cs[3] == 4 &&
          ks[3] == 1))
        return EC_E_UNSUPPORTED;
    typedef Cfg<32, 16, 512, 16, 0, 8, 0, 4, 0, 4, 1> G4;
    switch (stage + (states ? 4 : 0)) {
        case 0: return launch_fv<T, G4, 0>(ctx, n, l, v, p, A, tz, &rk);
        case 1: return launch_fv<T, G4, 1>(ctx, n, l, v, p, A, tz, &rk);
        case 2: return launch_fv<T, G4, 2>(ctx, n, l, v, p, A, tz, &rk);
        case 3: return launch_fv<T, G4, 3>(ctx, n, l, v, p, A, tz, &rk);
        case 4: return launch_fv<T, G4, 4>(ctx, n, l, v, p, A, tz, &rk);
        case 5: return launch_fv<T, G4, 5>(ctx, n, l, v, p, A, tz, &rk);
        case 6: return launch_fv<T, G4, 6>(ctx, n, l, v, p, A, tz, &rk);
        default: return launch_fv<T, G4, 7>(ctx, n, l, v, p, A, tz, &rk);
    }
}

template <typename T>
int filter_vel_t(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, int nsets, const ec_vel_set *sets,
                 int tz) {
    // the four-set volume/face configuration of the 1024^3 workload in one launch
    if (nsets == 4) {
        FArgs A;
        int cs[4], ks[4];
        for (int s = 0; s < 4; s++) {
            cs[s] = n / sets[s].nbar;
            ks[s] = sets[s].kind;
            for (int a = 0; a < 3; a++) A.s[s].out[a] = (double *)sets[s].out[a];
            A.s[s].nb = sets[s].nbar;
        }
        if (cs[0] == 16 && ks[0] == 0 && cs[1] == 8 && ks[1] == 0 && cs[2] == 4 && ks[2] == 0 && cs[3] == 4 &&
            ks[3] == 1) {
            const int rc = launch_fv<T, Cfg<32, 16, 512, 16, 0, 8, 0, 4, 0, 4, 1>>(ctx, n, l, v, p, A, tz);
            if (rc != EC_E_UNSUPPORTED) return rc;
        }
    }
    // otherwise one launch per set
    for (int s = 0; s < nsets; s++) {
        FArgs A;
        for (int a = 0; a < 3; a++) A.s[0].out[a] = (double *)sets[s].out[a];
        A.s[0].nb = sets[s].nbar;
        const int c = n / sets[s].nbar;
        int rc = sets[s].kind == 0 ? launch_single<T, 0>(ctx, n, l, v, p, A, c, tz)
                                   : launch_single<T, 1>(ctx, n, l, v, p, A, c, tz);
        if (rc) return rc;
    }
    return EC_OK;
}

}  // namespace

// tx is kept for the interface; the column tile edge is fixed per set list
extern "C" int ec_filter_vel_set_tiles(ec_ctx *ctx, int tx, int tz) {
    EC_REQUIRE(ctx, ctx && tx >= 1 && tz >= 1, "ec_filter_vel_set_tiles: bad argument");
    ctx->fv_tz = tz;
    return EC_OK;
}

extern "C" int ec_ns_filter_vel(ec_ctx *ctx, int dtype, int n, double l, void *const v[3], const void *p, int nsets,
                                const ec_vel_set *sets) {
    EC_REQUIRE(ctx, ctx && v && (nsets == 0 || sets), "ec_ns_filter_vel: null argument");
    for (int s = 0; s < nsets; s++) {
        EC_REQUIRE(ctx, sets[s].nbar >= 1 && n % sets[s].nbar == 0, "set %d: n=%d not a multiple of nbar=%d", s, n,
                   sets[s].nbar);
        EC_REQUIRE(ctx, sets[s].kind == 0 || sets[s].kind == 1, "set %d: kind %d", s, sets[s].kind);
        for (int a = 0; a < 3; a++) EC_REQUIRE(ctx, sets[s].out[a], "set %d: null output", s);
    }
    int rc = dtype == EC_F64 ? filter_vel_t<double>(ctx, n, l, v, p, nsets, sets, ctx->fv_tz)
                             : filter_vel_t<float>(ctx, n, l, v, p, nsets, sets, ctx->fv_tz);
    if (rc != EC_E_UNSUPPORTED) return rc;
    // any other compression: the per-quantity separable filter
    static const int src_k[3] = {EC_SRC_K1, EC_SRC_K2, EC_SRC_K3};
    for (int s = 0; s < nsets; s++) {
        const int c = n / sets[s].nbar;
        std::vector<std::vector<double>> keep;
        std::vector<ec_filter_quantity> q(3);
        for (int a = 0; a < 3; a++) {
            q[a].source = p ? src_k[a] : EC_SRC_U1 + a;
            for (int d = 0; d < 3; d++) {
                std::vector<double> w;
                int off;
                if (d != a) {
                    w.assign(c, 1.0 / c);
                    off = -(c - 1);
                } else if (sets[s].kind == 1) {
                    w.assign(1, 1.0);
                    off = 0;
                } else if (c & 1) {
                    w.assign(c, 1.0 / c);
                    off = -(c / 2);
                } else {
                    w.assign(c + 1, 1.0 / c);
                    w[0] = w[c] = 0.5 / c;
                    off = -(c / 2);
                }
                keep.push_back(w);
                q[a].len[d] = (int)w.size();
                q[a].off[d] = off;
            }
            q[a].out = sets[s].out[a];
        }
        for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) q[a].w[d] = keep[a * 3 + d].data();
        // velocity sources come from v (the filter kernel reads them as u)
        rc = ec_filter_coarse(ctx, dtype, n, sets[s].nbar, l, 0.0, v, p, 3, q.data());
        if (rc) return rc;
    }
    return EC_OK;
}

extern "C" int ec_ns_filter_vel_rk4(ec_ctx *ctx, int dtype, int n, double l, void *const du[3], const void *p,
                                    int nsets, const ec_vel_set *sets, int stage, double dt, void *const u0[3],
                                    void *const us[3], void *const acc[3]) {
    EC_REQUIRE(ctx, ctx && du && p && u0 && acc && (nsets == 0 || sets), "ec_ns_filter_vel_rk4: null argument");
    EC_REQUIRE(ctx, stage >= 0 && stage <= 3, "ec_ns_filter_vel_rk4: stage %d out of range", stage);
    EC_REQUIRE(ctx, stage == 3 || us, "ec_ns_filter_vel_rk4: stage state required before the last stage");
    for (int s = 0; s < nsets; s++) {
        EC_REQUIRE(ctx, sets[s].nbar >= 1 && n % sets[s].nbar == 0, "set %d: n=%d not a multiple of nbar=%d", s, n,
                   sets[s].nbar);
        EC_REQUIRE(ctx, sets[s].kind == 0 || sets[s].kind == 1, "set %d: kind %d", s, sets[s].kind);
        for (int a = 0; a < 3; a++) EC_REQUIRE(ctx, sets[s].out[a], "set %d: null output", s);
    }
    RkArgs rk{};
    for (int a = 0; a < 3; a++) {
        rk.u0[a] = u0[a];
        rk.us[a] = us ? us[a] : nullptr;
        rk.acc[a] = acc[a];
    }
    rk.dt = dt;
    int rc = dtype == EC_F64 ? filter_vel_rk4_t<double>(ctx, n, l, du, p, nsets, sets, ctx->fv_tz, stage, rk, false)
                             : filter_vel_rk4_t<float>(ctx, n, l, du, p, nsets, sets, ctx->fv_tz, stage, rk, false);
    if (rc != EC_E_UNSUPPORTED) return rc;
    // any other set list: the filter, then the stage update
    if ((rc = ec_ns_filter_vel(ctx, dtype, n, l, du, p, nsets, sets))) return rc;
    return ec_ns_rk4_update(ctx, dtype, n, l, stage, dt, u0, us, du, p, acc);
}

extern "C" int ec_ns_filter_vel_rk4_states(ec_ctx *ctx, int dtype, int n, double l, void *const du[3],
                                           const void *p, int nsets, const ec_vel_set *sets, int stage, double dt,
                                           void *const u0[3], void *const ust[9]) {
    EC_REQUIRE(ctx, ctx && du && p && u0 && ust && (nsets == 0 || sets), "ec_ns_filter_vel_rk4_states: null argument");
    EC_REQUIRE(ctx, stage >= 0 && stage <= 3, "ec_ns_filter_vel_rk4_states: stage %d out of range", stage);
    for (int q = 0; q < 9; q++)
        if (stage == 3 || q / 3 == stage)
            EC_REQUIRE(ctx, ust[q], "ec_ns_filter_vel_rk4_states: null stage state %d", q);
    for (int s = 0; s < nsets; s++) {
        EC_REQUIRE(ctx, sets[s].nbar >= 1 && n % sets[s].nbar == 0, "set %d: n=%d not a multiple of nbar=%d", s, n,
                   sets[s].nbar);
        EC_REQUIRE(ctx, sets[s].kind == 0 || sets[s].kind == 1, "set %d: kind %d", s, sets[s].kind);
        for (int a = 0; a < 3; a++) EC_REQUIRE(ctx, sets[s].out[a], "set %d: null output", s);
    }
    RkArgs rk{};
    for (int a = 0; a < 3; a++) {
        rk.u0[a] = u0[a];
        for (int s = 0; s < 3; s++) rk.ust[s][a] = ust[3 * s + a];
    }
    rk.dt = dt;
    int rc = dtype == EC_F64 ? filter_vel_rk4_t<double>(ctx, n, l, du, p, nsets, sets, ctx->fv_tz, stage, rk, true)
                             : filter_vel_rk4_t<float>(ctx, n, l, du, p, nsets, sets, ctx->fv_tz, stage, rk, true);
    if (rc != EC_E_UNSUPPORTED) return rc;
    if ((rc = ec_ns_filter_vel(ctx, dtype, n, l, du, p, nsets, sets))) return rc;
    return ec_ns_rk4_update_states(ctx, dtype, n, l, stage, dt, u0, ust, du, p);
}
