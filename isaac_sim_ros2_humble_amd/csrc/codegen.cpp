// Robot-specialised HIP source generation (replaces the NVRTC source generators
// cudampl/CUDAMPLib/include/spaces/KernelGenerator.h:222-522 and
// cudampl/CPRRTC/src/RobotSolver.cu:475-505, 801-1110).
//
// The FK is emitted as straight-line FP32 code: one SSA temporary per non-trivial operation,
// constants as exact hex-float literals.  A small symbolic evaluator folds constant
// sub-expressions on the host (fixed-joint chains, inactive joints at their default value) and
// drops terms whose coefficient is a constant 0 (and turns coefficient +-1 products into
// add/sub).  Both rewrites are exact in IEEE arithmetic (up to the sign of a zero), so the
// generated code evaluates the same FP32 rounding sequence as the dense rule set:
//   dot3(a,b) = fma(a0,b0, fma(a1,b1, a2*b2)),  row(a,b,t) = fma(a0,b0, fma(a1,b1, fma(a2,b2,t)))
//   fixed T = P o O;  revolute T = (P o O) o Rot(axis,q);  prismatic T = P o (O.R, O.R*(axis*q)+O.t)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <array>
#include <map>
#include <sstream>

#include "mp_internal.h"

namespace mp {

namespace {

// Same algorithm as mp_sincosf in the kernel source (Cody-Waite + Cephes polynomials); used
// to fold joints that sit at a constant default value.
void host_sincosf(float x, float& s_out, float& c_out) {
    const float two_over_pi = 0.636619772367581343f;
    float kf = std::rint(x * two_over_pi);
    int k = (int)kf;
    float r = std::fmaf(-kf, 1.5703125f, x);
    r = std::fmaf(-kf, 4.837512969970703125e-4f, r);
    r = std::fmaf(-kf, 7.54978995489188216e-8f, r);
    float z = r * r;
    float sp = std::fmaf(std::fmaf(std::fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
    float cp = std::fmaf(std::fmaf(std::fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f),
                         z * z, std::fmaf(-0.5f, z, 1.0f));
    int q = k & 3;
    float s0 = (q & 1) ? cp : sp;
    float c0 = (q & 1) ? sp : cp;
    s_out = (q & 2) ? -s0 : s0;
    c_out = ((q + 1) & 2) ? -c0 : c0;
}

std::string lit(float f) {
    if (f == 0.0f) return std::signbit(f) ? "(-0.0f)" : "0.0f";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%af", (double)f);
    std::string s(buf);
    if (s[0] == '-') return "(" + s + ")";
    return s;
}

struct V {
    bool k;
    float c;
    std::string e;
};
V K(float c) { return V{true, c, ""}; }

class Emitter {
  public:
    std::ostringstream os;
    int n = 0;
    std::string ind = "        ";
    std::string pfx = "t";  // temporary names: pfx + counter

    std::string ref(const V& v) const { return v.k ? lit(v.c) : v.e; }
    V tmp(const std::string& expr) {
        std::string name = pfx + std::to_string(n++);
        os << ind << "const float " << name << " = " << expr << ";\n";
        return V{false, 0.0f, name};
    }
    V neg(const V& a) {
        if (a.k) return K(-a.c);
        return tmp("-" + a.e);
    }
    V mul(V a, V b) {
        if (a.k && b.k) return K(a.c * b.c);
        if (a.k) std::swap(a, b);
        if (b.k) {
            if (b.c == 0.0f) return K(0.0f);
            if (b.c == 1.0f) return a;
            if (b.c == -1.0f) return neg(a);
        }
        return tmp(ref(a) + " * " + ref(b));
    }
    V add(const V& a, const V& b) {
        if (a.k && b.k) return K(a.c + b.c);
        if (a.k && a.c == 0.0f) return b;
        if (b.k && b.c == 0.0f) return a;
        return tmp(ref(a) + " + " + ref(b));
    }
    V sub(const V& a, const V& b) {
        if (a.k && b.k) return K(a.c - b.c);
        if (b.k && b.c == 0.0f) return a;
        if (a.k && a.c == 0.0f) return neg(b);
        return tmp(ref(a) + " - " + ref(b));
    }
    V fma(V a, V b, const V& c) {
        if ((a.k && a.c == 0.0f) || (b.k && b.c == 0.0f)) return c;
        if (a.k && b.k && c.k) return K(std::fmaf(a.c, b.c, c.c));
        if (b.k && !a.k) std::swap(a, b);
        if (a.k && a.c == 1.0f) return add(b, c);
        if (a.k && a.c == -1.0f) return sub(c, b);
        if (c.k && c.c == 0.0f) return mul(a, b);
        return tmp("__builtin_fmaf(" + ref(a) + ", " + ref(b) + ", " + ref(c) + ")");
    }
    V dot3(const V& a0, const V& b0, const V& a1, const V& b1, const V& a2, const V& b2) {
        return fma(a0, b0, fma(a1, b1, mul(a2, b2)));
    }
    V row(const V& a0, const V& b0, const V& a1, const V& b1, const V& a2, const V& b2, const V& t) {
        return fma(a0, b0, fma(a1, b1, fma(a2, b2, t)));
    }
};

struct Aff {
    V R[9];
    V t[3];
};

Aff identity() {
    Aff a;
    for (int i = 0; i < 9; i++) a.R[i] = K((i % 4 == 0) ? 1.0f : 0.0f);
    for (int i = 0; i < 3; i++) a.t[i] = K(0.0f);
    return a;
}

Aff compose(Emitter& E, const Aff& A, const Aff& B) {
    Aff C;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            C.R[3 * i + j] = E.dot3(A.R[3 * i], B.R[j], A.R[3 * i + 1], B.R[3 + j], A.R[3 * i + 2], B.R[6 + j]);
        C.t[i] = E.row(A.R[3 * i], B.t[0], A.R[3 * i + 1], B.t[1], A.R[3 * i + 2], B.t[2], A.t[i]);
    }
    return C;
}

double snap(double v) {
    if (std::fabs(v) < 1e-12) return 0.0;
    if (std::fabs(v - 1.0) < 1e-12) return 1.0;
    if (std::fabs(v + 1.0) < 1e-12) return -1.0;
    return v;
}

// axis mode: 0/1/2 = +-x/y/z exactly, 3 = generic, -1 = zero axis
int axis_mode(const RobotModel& m, int i, int& sign, float nf[3]) {
    double ax[3] = {snap(m.axis[3 * i]), snap(m.axis[3 * i + 1]), snap(m.axis[3 * i + 2])};
    int nz = 0, which = -1;
    for (int a = 0; a < 3; a++)
        if (ax[a] != 0.0) {
            nz++;
            which = a;
        }
    sign = 1;
    if (nz == 0) {
        for (int a = 0; a < 3; a++) nf[a] = 0.0f;
        return -1;
    }
    if (nz == 1 && std::fabs(ax[which]) == 1.0) {
        sign = ax[which] > 0 ? 1 : -1;
        for (int a = 0; a < 3; a++) nf[a] = (float)ax[a];
        return which;
    }
    if (m.jtype[i] == 1) {
        double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        for (int a = 0; a < 3; a++) nf[a] = (float)(ax[a] / n);
    } else {
        for (int a = 0; a < 3; a++) nf[a] = (float)ax[a];
    }
    return 3;
}

Aff rotation(Emitter& E, int mode, int sign, const float nf[3], const V& s, const V& c) {
    Aff R;
    for (int i = 0; i < 9; i++) R.R[i] = K(0.0f);
    for (int i = 0; i < 3; i++) R.t[i] = K(0.0f);
    V ss = sign > 0 ? s : E.neg(s);
    if (mode == 2) {
        R.R[0] = c; R.R[1] = E.neg(ss); R.R[3] = ss; R.R[4] = c; R.R[8] = K(1.0f);
    } else if (mode == 1) {
        R.R[0] = c; R.R[2] = ss; R.R[4] = K(1.0f); R.R[6] = E.neg(ss); R.R[8] = c;
    } else if (mode == 0) {
        R.R[0] = K(1.0f); R.R[4] = c; R.R[5] = E.neg(ss); R.R[7] = ss; R.R[8] = c;
    } else if (mode == 3) {
        V nx = K(nf[0]), ny = K(nf[1]), nz = K(nf[2]);
        V oc = E.sub(K(1.0f), c);
        R.R[0] = E.fma(E.mul(nx, nx), oc, c);
        R.R[1] = E.fma(E.mul(nx, ny), oc, E.neg(E.mul(nz, s)));
        R.R[2] = E.fma(E.mul(nx, nz), oc, E.mul(ny, s));
        R.R[3] = E.fma(E.mul(ny, nx), oc, E.mul(nz, s));
        R.R[4] = E.fma(E.mul(ny, ny), oc, c);
        R.R[5] = E.fma(E.mul(ny, nz), oc, E.neg(E.mul(nx, s)));
        R.R[6] = E.fma(E.mul(nz, nx), oc, E.neg(E.mul(ny, s)));
        R.R[7] = E.fma(E.mul(nz, ny), oc, E.mul(nx, s));
        R.R[8] = E.fma(E.mul(nz, nz), oc, c);
    } else {
        R.R[0] = R.R[4] = R.R[8] = K(1.0f);
    }
    return R;
}

std::string farr(const std::vector<float>& v) {
    std::string s;
    for (size_t i = 0; i < v.size(); i++) {
        if (i) s += ", ";
        s += lit(v[i]);
    }
    return s;
}

}  // namespace

float sqrt_threshold(float x) {
    // smallest float y >= 0 with sqrtf(y) >= x (sqrtf correctly rounded)
    float y = x * x;
    if (std::sqrt(y) >= x) {
        while (y > 0.0f && std::sqrt(std::nextafter(y, 0.0f)) >= x) y = std::nextafter(y, 0.0f);
    } else {
        while (std::sqrt(y) < x) y = std::nextafter(y, INFINITY);
    }
    return y;
}

namespace {

constexpr double kCullMargin = 1e-4;  // m, conservative slack on every bounding-sphere test

struct SphereLink {
    int link;
    std::vector<int> spheres;
    double c[3];  // bounding-sphere centre in the link frame
    double R;     // bounding-sphere radius (contains every sphere of the link)
};

std::vector<SphereLink> sphere_links(const RobotModel& m) {
    std::vector<SphereLink> out;
    std::vector<int> seen(m.L, -1);
    for (int s = 0; s < m.S; s++) {
        int l = m.sphere_link[s];
        if (seen[l] < 0) {
            seen[l] = (int)out.size();
            out.push_back(SphereLink{l, {}, {0, 0, 0}, 0.0});
        }
        out[seen[l]].spheres.push_back(s);
    }
    for (auto& sl : out) {
        double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
        for (int s : sl.spheres)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], (double)m.sphere_pos[3 * s + a] - m.sphere_r[s]);
                hi[a] = std::max(hi[a], (double)m.sphere_pos[3 * s + a] + m.sphere_r[s]);
            }
        for (int a = 0; a < 3; a++) sl.c[a] = (float)(0.5 * (lo[a] + hi[a]));
        double R = 0.0;
        for (int s : sl.spheres) {
            double d2 = 0;
            for (int a = 0; a < 3; a++) {
                double d = m.sphere_pos[3 * s + a] - sl.c[a];
                d2 += d * d;
            }
            R = std::max(R, std::sqrt(d2) + m.sphere_r[s]);
        }
        sl.R = R;
    }
    return out;
}

// flop costs of the planning cost model (FMA = 2, add/mul/min/max/cmp = 1)
constexpr int kXform = 18, kSelf = 9, kSph = 11, kBox = 34, kCyl = 35, kCap = 28;
constexpr int kCentreChunk = 16;  // spheres per staged chunk; the LDS row is MP_CENTRE_ROW floats

}  // namespace

// Link frames of every link (KernelGenerator.h:254-348 composition order), emitted into E with
// constant folding; returns the symbolic frames.
std::vector<Aff> emit_frames(Emitter& E, const RobotModel& m, int& fk_flops) {
    std::vector<Aff> T(m.L);
    T[0] = identity();
    int ntrig = 0;
    for (int i = 1; i < m.L; i++) {
        const Aff& P = T[m.parent[i]];
        Aff O;
        for (int a = 0; a < 3; a++) {
            for (int b = 0; b < 3; b++) O.R[3 * a + b] = K((float)snap(m.origin[16 * i + 4 * a + b]));
            O.t[a] = K((float)snap(m.origin[16 * i + 4 * a + 3]));
        }
        V v = m.active[i] ? V{false, 0.0f, "q[" + std::to_string(m.active_index[i]) + "]"} : K(m.defaults[i]);
        int sign = 1;
        float nf[3];
        int mode = axis_mode(m, i, sign, nf);
        if (m.jtype[i] == 1) {
            fk_flops += 130;
            Aff M = compose(E, P, O);
            V s, c;
            if (v.k) {
                float sv, cv;
                host_sincosf(v.c, sv, cv);
                s = K(sv);
                c = K(cv);
            } else {
                std::string sn = "s" + std::to_string(ntrig), cn = "c" + std::to_string(ntrig);
                ntrig++;
                E.os << E.ind << "float " << sn << ", " << cn << ";\n";
                E.os << E.ind << "mp_sincosf(" << v.e << ", " << sn << ", " << cn << ");\n";
                s = V{false, 0.0f, sn};
                c = V{false, 0.0f, cn};
            }
            Aff Rq = rotation(E, mode, sign, nf, s, c);
            T[i] = compose(E, M, Rq);
        } else if (m.jtype[i] == 2) {
            fk_flops += 87;
            Aff B;
            V tq[3] = {E.mul(K(nf[0]), v), E.mul(K(nf[1]), v), E.mul(K(nf[2]), v)};
            for (int a = 0; a < 9; a++) B.R[a] = O.R[a];
            for (int a = 0; a < 3; a++) B.t[a] = E.row(O.R[3 * a], tq[0], O.R[3 * a + 1], tq[1], O.R[3 * a + 2], tq[2], O.t[a]);
            T[i] = compose(E, P, B);
        } else if (m.jtype[i] == 5) {
            fk_flops += 63;
            T[i] = compose(E, P, O);
        } else {
            T[i] = P;
        }
    }
    return T;
}

std::string robot_struct_source(const RobotModel& m, const std::string& sname) {
    // ---------------------------------------------------------------- forward kinematics
    Emitter E;
    int fk_flops = 0;
    const std::vector<Aff> T = emit_frames(E, m, fk_flops);
    // link poses [L][16], row-major 4x4 with the last row (0, 0, 0, 1)
    std::string poses_body;
    {
        Emitter EP;
        int unused = 0;
        const std::vector<Aff> TP = emit_frames(EP, m, unused);
        for (int i = 0; i < m.L; i++) {
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < 3; b++)
                    EP.os << EP.ind << "P[" << 16 * i + 4 * a + b << "] = " << EP.ref(TP[i].R[3 * a + b]) << ";\n";
                EP.os << EP.ind << "P[" << 16 * i + 4 * a + 3 << "] = " << EP.ref(TP[i].t[a]) << ";\n";
            }
            EP.os << EP.ind << "P[" << 16 * i + 12 << "] = 0.0f; P[" << 16 * i + 13 << "] = 0.0f; P[" << 16 * i + 14
                  << "] = 0.0f; P[" << 16 * i + 15 << "] = 1.0f;\n";
        }
        poses_body = EP.os.str();
    }
    const std::vector<SphereLink> SL = sphere_links(m);
    const int NL = (int)SL.size();
    fk_flops += kXform * NL;
    // F layout per sphere link k: [k*15 + 0..8] rotation (row major), [+9..11] translation,
    // [+12..14] world bounding-sphere centre.  Constant entries are returned as constants.
    std::vector<std::vector<V>> FV(NL, std::vector<V>(15));
    for (int k = 0; k < NL; k++) {
        const Aff& M = T[SL[k].link];
        V b[3];
        for (int a = 0; a < 3; a++)
            b[a] = E.row(M.R[3 * a], K((float)SL[k].c[0]), M.R[3 * a + 1], K((float)SL[k].c[1]), M.R[3 * a + 2],
                         K((float)SL[k].c[2]), M.t[a]);
        for (int e = 0; e < 9; e++) FV[k][e] = M.R[e];
        for (int a = 0; a < 3; a++) FV[k][9 + a] = M.t[a];
        for (int a = 0; a < 3; a++) FV[k][12 + a] = b[a];
        for (int e = 0; e < 15; e++) {
            if (!FV[k][e].k) {
                E.os << E.ind << "F[" << k * 15 + e << "] = " << FV[k][e].e << ";\n";
                FV[k][e] = V{false, 0.0f, "F[" + std::to_string(k * 15 + e) + "]"};
            }
        }
    }
    std::string fk_body = E.os.str();

    // sphere -> (sphere link index)
    std::vector<int> slk(m.S, -1);
    for (int k = 0; k < NL; k++)
        for (int s : SL[k].spheres) slk[s] = k;

    auto centre = [&](Emitter& X, int s, V out[3]) {
        const std::vector<V>& f = FV[slk[s]];
        V p[3] = {K(m.sphere_pos[3 * s]), K(m.sphere_pos[3 * s + 1]), K(m.sphere_pos[3 * s + 2])};
        for (int a = 0; a < 3; a++) out[a] = X.row(f[3 * a], p[0], f[3 * a + 1], p[1], f[3 * a + 2], p[2], f[9 + a]);
    };
    auto named = [&](Emitter& X, const V& v) -> std::string { return X.ref(v); };

    std::vector<float> rad(m.sphere_r.begin(), m.sphere_r.end()), boxte(m.S);
    for (int s = 0; s < m.S; s++) boxte[s] = std::min(m.sphere_r[s] * m.sphere_r[s], sqrt_threshold(m.sphere_r[s]));

    // ---------------------------------------------------------------- centres output
    Emitter C;
    for (int s = 0; s < m.S; s++) {
        V c[3];
        centre(C, s, c);
        C.os << C.ind << "out[" << 3 * s << "] = " << named(C, c[0]) << "; out[" << 3 * s + 1 << "] = " << named(C, c[1])
             << "; out[" << 3 * s + 2 << "] = " << named(C, c[2]) << ";\n";
    }
    std::string centres_body = C.os.str();

    // ---------------------------------------------------------------- staged centres output
    // The wave's centres [64][S][3] are one contiguous run of memory.  Written lane by lane, every
    // store instruction touches 64 lines 12 S bytes apart, and lines leave the cache before they
    // are complete.  Staged: chunks of kCentreChunk spheres go to the wave's LDS tile (one row of
    // 3 kCentreChunk + 1 floats per lane), and the wave copies each chunk out with consecutive
    // lanes on consecutive floats.  The values are the same expressions as write_centres.
    std::ostringstream staged;
    for (int s0 = 0; s0 < m.S; s0 += kCentreChunk) {
        const int s1 = std::min(m.S, s0 + kCentreChunk), cc = 3 * (s1 - s0);
        Emitter X;
        X.pfx = "c" + std::to_string(s0) + "_";
        X.ind = "            ";
        X.os << "        {\n";
        for (int s = s0; s < s1; s++) {
            V c[3];
            centre(X, s, c);
            for (int a = 0; a < 3; a++) X.os << X.ind << "row[" << 3 * (s - s0) + a << "] = " << named(X, c[a]) << ";\n";
        }
        X.os << "        }\n";
        X.os << "        mp_wave_sync();\n";
        X.os << "        for (int t = lane; t < live * " << cc << "; t += MP_WAVE) {\n"
             << "            const int L = t / " << cc << ", j = t - L * " << cc << ";\n"
             << "            out[(long long)L * " << 3 * m.S << " + " << 3 * s0 << " + j] = tile[L * MP_CENTRE_ROW + j];\n"
             << "        }\n"
             << "        mp_wave_sync();\n";
        staged << X.os.str();
    }
    std::string staged_body = staged.str();

    // ---------------------------------------------------------------- self collision blocks
    // link pairs in the order of the first sphere index of each link (the oracle's order)
    struct LP {
        int a, b;
        std::vector<std::pair<int, int>> pairs;
    };
    std::vector<LP> lps;
    for (int ka = 0; ka < NL; ka++)
        for (int kb = ka + 1; kb < NL; kb++) {
            int la = SL[ka].link, lb = SL[kb].link;
            if (!m.enabled[la * m.L + lb]) continue;
            LP lp{ka, kb, {}};
            for (int i : SL[ka].spheres)
                for (int j : SL[kb].spheres) lp.pairs.emplace_back(std::min(i, j), std::max(i, j));
            lps.push_back(lp);
        }
    auto pair_thr = [&](int i, int j, float& thr, float& sr) {
        float r = m.sphere_r[i] + m.sphere_r[j];
        thr = r * r;
        sr = std::sqrt(thr);
    };
    auto self_body = [&](bool cost) {
        Emitter X;
        int blk = 0;
        for (const LP& lp : lps) {
            const SphereLink& A = SL[lp.a];
            const SphereLink& B = SL[lp.b];
            float cull = (float)((A.R + B.R + kCullMargin) * (A.R + B.R + kCullMargin));
            // in a team, wave w evaluates the blocks blk with blk % W == w
            if (!cost) X.os << "        if (T.mine(" << blk << ")) {\n";
            blk++;
            X.os << "        {\n";
            V d[3];
            for (int a = 0; a < 3; a++) d[a] = X.sub(FV[lp.a][12 + a], FV[lp.b][12 + a]);
            V d2 = X.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
            X.os << "        work += " << kSelf << "u;\n";
            X.os << "        if (__ballot(" << X.ref(d2) << " < " << lit(cull) << ")) {\n";
            int nsp = (int)(A.spheres.size() + B.spheres.size());
            X.os << "            work += " << (int)lp.pairs.size() * kSelf + nsp * kXform << "u;\n";
            if (!cost) {
                // verdict: hold the centres of the link with fewer spheres and stream the other
                // link's spheres one at a time (the pair set is every sphere of A with every
                // sphere of B; OR-ing the contacts does not depend on their order)
                const bool a_small = A.spheres.size() <= B.spheres.size();
                const SphereLink& Hk = a_small ? A : B;
                const SphereLink& Sk = a_small ? B : A;
                std::map<int, std::array<V, 3>> held;
                for (int s : Hk.spheres) { V c[3]; centre(X, s, c); held[s] = {c[0], c[1], c[2]}; }
                for (int s : Sk.spheres) {
                    V c[3];
                    centre(X, s, c);
                    for (int hsph : Hk.spheres) {
                        float thr, sr;
                        pair_thr(std::min(hsph, s), std::max(hsph, s), thr, sr);
                        V e[3];
                        const int i = std::min(hsph, s);
                        const std::array<V, 3>& ci = i == s ? std::array<V, 3>{c[0], c[1], c[2]} : held[hsph];
                        const std::array<V, 3>& cj = i == s ? held[hsph] : std::array<V, 3>{c[0], c[1], c[2]};
                        for (int a = 0; a < 3; a++) e[a] = X.sub(ci[a], cj[a]);
                        V dd = X.dot3(e[0], e[0], e[1], e[1], e[2], e[2]);
                        X.os << "            h |= " << X.ref(dd) << " < " << lit(std::min(thr, sqrt_threshold(sr))) << ";\n";
                    }
                }
                X.os << "        }\n        }\n";
                X.os << "        if (mp_stop<MODE>(h, mask)) return h;\n        }\n";
                continue;
            }
            std::map<int, std::array<V, 3>> cs;
            for (int s : A.spheres) { V c[3]; centre(X, s, c); cs[s] = {c[0], c[1], c[2]}; }
            for (int s : B.spheres) { V c[3]; centre(X, s, c); cs[s] = {c[0], c[1], c[2]}; }
            for (auto& pr : lp.pairs) {
                int i = pr.first, j = pr.second;
                float thr, sr;
                pair_thr(i, j, thr, sr);
                V e[3];
                for (int a = 0; a < 3; a++) e[a] = X.sub(cs[i][a], cs[j][a]);
                V dd = X.dot3(e[0], e[0], e[1], e[1], e[2], e[2]);
                if (cost)
                    X.os << "            part += " << X.ref(dd) << " < " << lit(thr) << " ? " << lit(sr) << " - __builtin_sqrtf("
                         << X.ref(dd) << ") : 0.0f;\n";
                else
                    X.os << "            h |= " << X.ref(dd) << " < " << lit(std::min(thr, sqrt_threshold(sr))) << ";\n";
            }
            X.os << "        }\n        }\n";
            if (!cost) X.os << "        if (mp_stop<MODE>(h, mask)) return h;\n        }\n";
        }
        return X.os.str();
    };

    // ---------------------------------------------------------------- link groups
    // Obstacle-level culling: the sphere links are split into contiguous groups; per lane, a
    // group's bounding sphere is centred on its middle link's bounding-sphere centre with radius
    // max(|b_k - c| + R_k) + 2 margins.  An obstacle whose bounding sphere misses every lane's
    // group sphere cannot reach any link of the group (triangle inequality), so the group's link
    // tests are skipped.
    std::vector<std::vector<int>> groups;
    {
        const int G = NL <= 2 ? 1 : std::min(4, (NL + 2) / 3);
        for (int g = 0; g < G; g++) {
            std::vector<int> v;
            for (int k = g * NL / G; k < (g + 1) * NL / G; k++) v.push_back(k);
            if (!v.empty()) groups.push_back(v);
        }
    }
    Emitter GP;
    GP.pfx = "gp";
    std::vector<std::array<V, 4>> gsph;
    for (const auto& g : groups) {
        const int km = g[g.size() / 2];
        V c[3] = {FV[km][12], FV[km][13], FV[km][14]};
        V r = K(0.0f);
        bool first = true;
        for (int k : g) {
            const float Rk = (float)(SL[k].R + 2.0 * kCullMargin);
            V term;
            if (k == km) {
                term = K(Rk);
            } else {
                V d[3] = {GP.sub(FV[k][12], c[0]), GP.sub(FV[k][13], c[1]), GP.sub(FV[k][14], c[2])};
                V dd = GP.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
                V sq = dd.k ? K(std::sqrt(dd.c)) : GP.tmp("__builtin_sqrtf(" + GP.ref(dd) + ")");
                term = GP.add(sq, K(Rk));
            }
            if (first) {
                r = term;
                first = false;
            } else if (r.k && term.k) {
                r = K(std::max(r.c, term.c));
            } else {
                r = GP.tmp("fmaxf(" + GP.ref(r) + ", " + GP.ref(term) + ")");
            }
        }
        gsph.push_back({c[0], c[1], c[2], r});
    }
    const std::string group_prologue = GP.os.str();

    // ---------------------------------------------------------------- environment blocks
    // kind: 0 box, 1 sphere, 2 capsule, 3 cylinder.  The obstacle parameters are named
    // m0..m11 lo{x,y,z} hi{x,y,z} (box), ox oy oz orad (sphere), x1 y1 z1 xv yv zv rdv crad
    // (capsule), m0..m11 crad hh (cylinder) in the generated loops.
    auto env_block = [&](int kind, bool cost, const char* detail) {
        Emitter X;
        V mm[12];
        for (int i = 0; i < 12; i++) mm[i] = V{false, 0.0f, "m" + std::to_string(i)};
        // obstacle bounding sphere (centre, radius): the sphere itself, or the record's bound
        const V oc[3] = {V{false, 0.0f, kind == 1 ? "ox" : "ocx"}, V{false, 0.0f, kind == 1 ? "oy" : "ocy"},
                         V{false, 0.0f, kind == 1 ? "oz" : "ocz"}};
        const std::string orr = kind == 1 ? "orad" : "orr";
        const bool cull_group = groups.size() > 1;
        // all group tests of this obstacle first (straight-line code, no branch in between), then
        // one branch for the common case that no lane is near any group
        if (cull_group) {
            std::string any;
            for (size_t gi = 0; gi < groups.size(); gi++) {
                V gd[3] = {X.sub(gsph[gi][0], oc[0]), X.sub(gsph[gi][1], oc[1]), X.sub(gsph[gi][2], oc[2])};
                V gd2 = X.dot3(gd[0], gd[0], gd[1], gd[1], gd[2], gd[2]);
                V gs = X.tmp(X.ref(gsph[gi][3]) + " + " + orr);
                X.os << "            const u64 gb" << gi << " = __ballot(" << X.ref(gd2) << " < " << X.ref(gs) << " * "
                     << X.ref(gs) << ");\n";
                any += (gi ? " | gb" : "gb") + std::to_string(gi);
            }
            X.os << "            work += " << kSph * (int)groups.size() << "u;\n";
            X.os << "            if (" << any << ") {\n" << detail;
        } else {
            X.os << detail;
        }
        for (size_t gi = 0; gi < groups.size(); gi++) {
        if (cull_group) X.os << "            {\n            if (gb" << gi << ") {\n";
        for (int k : groups[gi]) {
            const SphereLink& L = SL[k];
            float Rm = (float)(L.R + kCullMargin);
            int n = (int)L.spheres.size();
            X.os << "            {\n";
            V b[3] = {FV[k][12], FV[k][13], FV[k][14]};
            std::string cond;
            int tcost = kind == 0 ? kBox : kind == 1 ? kSph : kind == 2 ? kCap : kCyl;
            if (kind == 0 || kind == 3) {
                V l[3];
                for (int a = 0; a < 3; a++) l[a] = X.row(mm[4 * a], b[0], mm[4 * a + 1], b[1], mm[4 * a + 2], b[2], mm[4 * a + 3]);
                if (kind == 0) {
                    const char* lo[3] = {"lox", "loy", "loz"};
                    const char* hi[3] = {"hix", "hiy", "hiz"};
                    V d[3];
                    for (int a = 0; a < 3; a++)
                        d[a] = X.tmp(X.ref(l[a]) + " - fminf(fmaxf(" + X.ref(l[a]) + ", " + lo[a] + "), " + hi[a] + ")");
                    V d2 = X.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
                    cond = X.ref(d2) + " < " + lit(Rm * Rm);
                } else {
                    V sdf = X.tmp("mp_cyl_sdf(" + X.ref(l[0]) + ", " + X.ref(l[1]) + ", " + X.ref(l[2]) + ", crad, hh)");
                    cond = X.ref(sdf) + " < " + lit(Rm);
                }
            } else if (kind == 1) {
                V d[3] = {X.sub(b[0], V{false, 0, "ox"}), X.sub(b[1], V{false, 0, "oy"}), X.sub(b[2], V{false, 0, "oz"})};
                V d2 = X.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
                V sr = X.tmp(lit(Rm) + " + orad");
                cond = X.ref(d2) + " < " + X.ref(sr) + " * " + X.ref(sr);
            } else {
                V d2 = X.tmp("mp_cap_d2(" + X.ref(b[0]) + ", " + X.ref(b[1]) + ", " + X.ref(b[2]) + ", x1, y1, z1, xv, yv, zv, rdv)");
                V sr = X.tmp(lit(Rm) + " + crad");
                cond = X.ref(d2) + " < " + X.ref(sr) + " * " + X.ref(sr);
            }
            X.os << "            work += " << tcost << "u;\n";
            X.os << "            if (__ballot(" << cond << ")) {\n";
            X.os << "                work += " << n * (kXform + tcost) << "u;\n";
            // spheres and capsules: contact <=> d2 < (r+R)^2 and sqrtf(d2) < r+R (the cost
            // (r+R) - sqrtf(d2) is then > 0); a fast test with a 1e-6 relative margin, then the
            // exact test for lanes in the thin band below the threshold
            const bool banded = (kind == 1 || kind == 2) && !cost;
            if (banded) X.os << "                bool band = false;\n";
            std::vector<std::string> d2n, srn;
            for (int s : L.spheres) {
                V c[3];
                centre(X, s, c);
                float r = m.sphere_r[s];
                if (kind == 0) {
                    V l[3];
                    for (int a = 0; a < 3; a++) l[a] = X.row(mm[4 * a], c[0], mm[4 * a + 1], c[1], mm[4 * a + 2], c[2], mm[4 * a + 3]);
                    const char* lo[3] = {"lox", "loy", "loz"};
                    const char* hi[3] = {"hix", "hiy", "hiz"};
                    V d[3];
                    for (int a = 0; a < 3; a++)
                        d[a] = X.tmp(X.ref(l[a]) + " - fminf(fmaxf(" + X.ref(l[a]) + ", " + lo[a] + "), " + hi[a] + ")");
                    V d2 = X.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
                    if (cost)
                        X.os << "                part += " << X.ref(d2) << " < " << lit(r * r) << " ? " << lit(r)
                             << " - __builtin_sqrtf(" << X.ref(d2) << ") : 0.0f;\n";
                    else
                        X.os << "                hit |= " << X.ref(d2) << " < " << lit(boxte[s]) << ";\n";
                } else if (kind == 1) {
                    V d[3] = {X.sub(c[0], V{false, 0, "ox"}), X.sub(c[1], V{false, 0, "oy"}), X.sub(c[2], V{false, 0, "oz"})};
                    V d2 = X.dot3(d[0], d[0], d[1], d[1], d[2], d[2]);
                    V sr = X.tmp(lit(r) + " + orad");
                    if (cost) {
                        X.os << "                part += " << X.ref(d2) << " < " << X.ref(sr) << " * " << X.ref(sr)
                             << " ? fmaxf(0.0f, " << X.ref(sr) << " - __builtin_sqrtf(" << X.ref(d2) << ")) : 0.0f;\n";
                    } else {
                        V T2 = X.tmp(X.ref(sr) + " * " + X.ref(sr));
                        X.os << "                hit |= " << X.ref(d2) << " < " << X.ref(T2) << " * 0.999999f;\n";
                        X.os << "                band |= " << X.ref(d2) << " < " << X.ref(T2) << ";\n";
                        d2n.push_back(X.ref(d2));
                        srn.push_back(X.ref(sr));
                    }
                } else if (kind == 2) {
                    V d2 = X.tmp("mp_cap_d2(" + X.ref(c[0]) + ", " + X.ref(c[1]) + ", " + X.ref(c[2]) + ", x1, y1, z1, xv, yv, zv, rdv)");
                    V rs = X.tmp(lit(r) + " + crad");
                    if (cost) {
                        X.os << "                part += " << X.ref(d2) << " < " << X.ref(rs) << " * " << X.ref(rs)
                             << " ? fmaxf(0.0f, " << X.ref(rs) << " - __builtin_sqrtf(" << X.ref(d2) << ")) : 0.0f;\n";
                    } else {
                        V T2 = X.tmp(X.ref(rs) + " * " + X.ref(rs));
                        X.os << "                hit |= " << X.ref(d2) << " < " << X.ref(T2) << " * 0.999999f;\n";
                        X.os << "                band |= " << X.ref(d2) << " < " << X.ref(T2) << ";\n";
                        d2n.push_back(X.ref(d2));
                        srn.push_back(X.ref(rs));
                    }
                } else {
                    V l[3];
                    for (int a = 0; a < 3; a++) l[a] = X.row(mm[4 * a], c[0], mm[4 * a + 1], c[1], mm[4 * a + 2], c[2], mm[4 * a + 3]);
                    V sdf = X.tmp("mp_cyl_sdf(" + X.ref(l[0]) + ", " + X.ref(l[1]) + ", " + X.ref(l[2]) + ", crad, hh)");
                    if (cost)
                        X.os << "                { const float pen = " << lit(r) << " - " << X.ref(sdf)
                             << "; part += pen > 0.0f ? pen : 0.0f; }\n";
                    else
                        X.os << "                hit |= " << X.ref(sdf) << " < " << lit(r) << ";\n";
                }
            }
            if (banded) {
                X.os << "                if (__ballot(band && !hit)) {\n";
                for (size_t t = 0; t < d2n.size(); t++)
                    X.os << "                    hit |= (" << d2n[t] << " < " << srn[t] << " * " << srn[t] << ") && (__builtin_sqrtf("
                         << d2n[t] << ") < " << srn[t] << ");\n";
                X.os << "                }\n";
            }
            X.os << "            }\n            }\n";
        }
        if (cull_group) X.os << "            }\n            }\n";
        }
        if (cull_group) X.os << "            }\n";
        return X.os.str();
    };

    const int P = (int)m.pi.size();
    std::ostringstream o;
    o << "struct " << sname << " {\n";
    o << "    static constexpr int DOF = " << m.dof << ";\n";
    o << "    static constexpr int S = " << m.S << ";\n";
    o << "    static constexpr int P = " << P << ";\n";
    o << "    static constexpr int NL = " << NL << ";\n";
    o << "    static constexpr int NF = " << NL * 15 << ";\n";
    o << "    static constexpr unsigned FK_WORK = " << fk_flops << "u;\n";
    o << "    static __device__ __forceinline__ void fk(const float* q, float* F) {\n" << fk_body << "    }\n";
    o << "    static constexpr int NLINK = " << m.L << ";\n";
    o << "    static __device__ __forceinline__ void poses(const float* q, float* P) {\n" << poses_body << "    }\n";
    o << "    static __device__ __forceinline__ void write_centres(const float* F, float* out) {\n" << centres_body << "    }\n";
    o << "    // out: the wave's first configuration; live: its configurations below n; tile: the wave's\n"
         "    // LDS tile of MP_WAVE x MP_CENTRE_ROW floats\n"
         "    static __device__ __forceinline__ void write_centres_staged(const float* F, float* out, int live, float* tile) {\n"
         "        const int lane = threadIdx.x & (MP_WAVE - 1);\n"
         "        float* row = tile + lane * MP_CENTRE_ROW;\n"
      << staged_body << "    }\n";
    o << "    template <int MODE>\n    static __device__ __forceinline__ bool self_hit(const float* F, bool h, u64 mask, unsigned& work, const Team& T) {\n"
      << self_body(false) << "        return h;\n    }\n";
    o << "    static __device__ __forceinline__ float self_cost(const float* F, unsigned& work) {\n        float part = 0.0f;\n"
      << self_body(true) << "        return part;\n    }\n";
    // an obstacle's bounding sphere first (its own section of the scene), its record only when a
    // link group is near it
    const char* box_load =
        "            const float* bb = sc.bbox + o * 4;\n"
        "            const float ocx = mp_sld(bb + 0), ocy = mp_sld(bb + 1), ocz = mp_sld(bb + 2), orr = mp_sld(bb + 3);\n";
    const char* box_detail =
        "            const float* b = sc.box + o * MP_BOX_W;\n"
        "            const float m0 = mp_sld(b + 0), m1 = mp_sld(b + 1), m2 = mp_sld(b + 2), m3 = mp_sld(b + 3);\n"
        "            const float m4 = mp_sld(b + 4), m5 = mp_sld(b + 5), m6 = mp_sld(b + 6), m7 = mp_sld(b + 7);\n"
        "            const float m8 = mp_sld(b + 8), m9 = mp_sld(b + 9), m10 = mp_sld(b + 10), m11 = mp_sld(b + 11);\n"
        "            const float lox = mp_sld(b + 12), loy = mp_sld(b + 13), loz = mp_sld(b + 14);\n"
        "            const float hix = mp_sld(b + 15), hiy = mp_sld(b + 16), hiz = mp_sld(b + 17);\n";
    const char* sph_load =
        "            const float* p = sc.sph + o * MP_SPH_W;\n"
        "            const float ox = mp_sld(p + 0), oy = mp_sld(p + 1), oz = mp_sld(p + 2), orad = mp_sld(p + 3);\n";
    const char* cap_load =
        "            const float* kb = sc.bcap + o * 4;\n"
        "            const float ocx = mp_sld(kb + 0), ocy = mp_sld(kb + 1), ocz = mp_sld(kb + 2), orr = mp_sld(kb + 3);\n";
    const char* cap_detail =
        "            const float* k = sc.cap + o * MP_CAP_W;\n"
        "            const float x1 = mp_sld(k + 0), y1 = mp_sld(k + 1), z1 = mp_sld(k + 2);\n"
        "            const float xv = mp_sld(k + 3), yv = mp_sld(k + 4), zv = mp_sld(k + 5);\n"
        "            const float rdv = mp_sld(k + 6), crad = mp_sld(k + 7);\n";
    const char* cyl_load =
        "            const float* cb = sc.bcyl + o * 4;\n"
        "            const float ocx = mp_sld(cb + 0), ocy = mp_sld(cb + 1), ocz = mp_sld(cb + 2), orr = mp_sld(cb + 3);\n";
    const char* cyl_detail =
        "            const float* b = sc.cyl + o * MP_CYL_W;\n"
        "            const float m0 = mp_sld(b + 0), m1 = mp_sld(b + 1), m2 = mp_sld(b + 2), m3 = mp_sld(b + 3);\n"
        "            const float m4 = mp_sld(b + 4), m5 = mp_sld(b + 5), m6 = mp_sld(b + 6), m7 = mp_sld(b + 7);\n"
        "            const float m8 = mp_sld(b + 8), m9 = mp_sld(b + 9), m10 = mp_sld(b + 10), m11 = mp_sld(b + 11);\n"
        "            const float crad = mp_sld(b + 12), hh = mp_sld(b + 13);\n";
    // early-out order: boxes, spheres, capsules, cylinders (cheapest per pair first)
    // in a team, wave w takes the obstacles o with o % W == w of every kind; the team's early-exit
    // flag (an LDS read) is looked at every fourth obstacle
    o << "    template <int MODE>\n    static __device__ __forceinline__ bool env_hit(const float* F, const SceneDev& sc, bool hit, u64 mask, unsigned& work, const Team& T) {\n";
    o << "        const int o0 = mp_uniform(T.w), os = mp_uniform(T.W);\n" << group_prologue;
    o << "        for (int o = o0, it = 1; o < mp_uniform(sc.nb); o += os, ++it) {\n" << box_load << env_block(0, false, box_detail)
      << "            if (mp_stop<MODE>(hit, mask) || ((it & 3) == 0 && T.stopped())) return hit;\n        }\n";
    o << "        for (int o = o0, it = 1; o < mp_uniform(sc.ns); o += os, ++it) {\n" << sph_load << env_block(1, false, "")
      << "            if (mp_stop<MODE>(hit, mask) || ((it & 3) == 0 && T.stopped())) return hit;\n        }\n";
    o << "        for (int o = o0, it = 1; o < mp_uniform(sc.ncap); o += os, ++it) {\n" << cap_load << env_block(2, false, cap_detail)
      << "            if (mp_stop<MODE>(hit, mask) || ((it & 3) == 0 && T.stopped())) return hit;\n        }\n";
    o << "        for (int o = o0, it = 1; o < mp_uniform(sc.nc); o += os, ++it) {\n" << cyl_load << env_block(3, false, cyl_detail)
      << "            if (mp_stop<MODE>(hit, mask) || ((it & 3) == 0 && T.stopped())) return hit;\n        }\n";
    o << "        return hit;\n    }\n";
    // full-evaluation costs in the oracle order: spheres, boxes, cylinders, capsules; culled
    // blocks contribute exact zeros, so skipping them leaves every partial sum unchanged
    // (per kind into pk[0..3]: the per-constraint costs of the sphere, cuboid, cylinder and
    // capsule constraints, States.h:205's [constraint][state] rows)
    o << "    static __device__ __forceinline__ void env_cost(const float* F, const SceneDev& sc, unsigned& work, float* pk) {\n";
    o << "        float part = 0.0f;\n" << group_prologue;
    o << "        for (int o = 0; o < mp_uniform(sc.ns); ++o) {\n" << sph_load << env_block(1, true, "") << "        }\n";
    o << "        pk[0] = part; part = 0.0f;\n";
    o << "        for (int o = 0; o < mp_uniform(sc.nb); ++o) {\n" << box_load << env_block(0, true, box_detail) << "        }\n";
    o << "        pk[1] = part; part = 0.0f;\n";
    o << "        for (int o = 0; o < mp_uniform(sc.nc); ++o) {\n" << cyl_load << env_block(3, true, cyl_detail) << "        }\n";
    o << "        pk[2] = part; part = 0.0f;\n";
    o << "        for (int o = 0; o < mp_uniform(sc.ncap); ++o) {\n" << cap_load << env_block(2, true, cap_detail) << "        }\n";
    o << "        pk[3] = part;\n    }\n";
    // joint bounds: BoundaryConstraint.cu:56-92, tolerance 1e-4 on the squared sum
    o << "    static __device__ __forceinline__ float bounds_acc(const float* q) {\n        float acc = 0.0f;\n";
    for (int j = 0; j < m.dof; j++) {
        o << "        { const float v = q[" << j << "]; const float d = v < " << lit(m.act_lo[j]) << " ? " << lit(m.act_lo[j])
          << " - v : (v > " << lit(m.act_hi[j]) << " ? v - " << lit(m.act_hi[j]) << " : 0.0f); acc = __builtin_fmaf(d, d, acc); }\n";
    }
    o << "        return acc;\n    }\n";
    o << "    static __device__ __forceinline__ bool bounds_ok(const float* q) { return bounds_acc(q) < 1e-4f; }\n";
    o << "    static __device__ __forceinline__ float bounds_cost(const float* q) { const float a = bounds_acc(q); "
         "return a < 1e-4f ? 0.0f : __builtin_sqrtf(a); }\n";
    // Philox sample in bounds
    o << "    static __device__ __forceinline__ void sample(u64 seed, u64 idx, float* q) {\n";
    for (int blk = 0; blk * 4 < m.dof; blk++) {
        o << "        { u32 c0 = (u32)idx, c1 = (u32)(idx >> 32), c2 = " << blk << "u, c3 = 0x6d70u;\n";
        o << "          mp_philox(c0, c1, c2, c3, (u32)seed, (u32)(seed >> 32));\n";
        const char* cn[4] = {"c0", "c1", "c2", "c3"};
        for (int w = 0; w < 4 && blk * 4 + w < m.dof; w++) {
            int j = blk * 4 + w;
            float range = m.act_hi[j] - m.act_lo[j];
            o << "          q[" << j << "] = __builtin_fmaf((float)(" << cn[w] << " >> 8) * 5.9604644775390625e-08f, " << lit(range)
              << ", " << lit(m.act_lo[j]) << ");\n";
        }
        o << "        }\n";
    }
    o << "    }\n";
    // Halton sample: point idx + 1 of the Halton sequence in the prime bases 3, 5, 7, ... (one
    // per active joint), rotated modulo 1 per seed (Cranley-Patterson) by Philox words drawn at
    // counter (0, 0, block, 0x6861), then scaled into the bounds as above
    static const unsigned kHaltonBases[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};
    o << "    static __device__ __forceinline__ void sample_halton(u64 seed, u64 idx, float* q) {\n";
    o << "        const u32 i = (u32)(idx + 1);\n";
    for (int blk = 0; blk * 4 < m.dof; blk++) {
        o << "        { u32 c0 = 0u, c1 = 0u, c2 = " << blk << "u, c3 = 0x6861u;\n";
        o << "          mp_philox(c0, c1, c2, c3, (u32)seed, (u32)(seed >> 32));\n";
        const char* cn[4] = {"c0", "c1", "c2", "c3"};
        for (int w = 0; w < 4 && blk * 4 + w < m.dof; w++) {
            int j = blk * 4 + w;
            float range = m.act_hi[j] - m.act_lo[j];
            o << "          { float v = mp_radical_inverse<" << kHaltonBases[j] << "u>(i) + (float)(" << cn[w]
              << " >> 8) * 5.9604644775390625e-08f; v = v >= 1.0f ? v - 1.0f : v;\n";
            o << "            q[" << j << "] = __builtin_fmaf(v, " << lit(range) << ", " << lit(m.act_lo[j]) << "); }\n";
        }
        o << "        }\n";
    }
    o << "    }\n";
    o << "};\n";
    return o.str();
}

std::string module_source(const std::vector<const RobotModel*>& robots) {
    std::ostringstream o;
    o << kernel_template_prelude();
    for (size_t k = 0; k < robots.size(); k++) o << robot_struct_source(*robots[k], "Robot" + std::to_string(k));
    o << kernel_template_body();
    for (size_t k = 0; k < robots.size(); k++) {
        std::string R = "Robot" + std::to_string(k), s = std::to_string(k);
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_fk_" << s
          << "(const float* q, long long n, float* c) { mp_fk_kernel<" << R << ">(q, n, c); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_poses_" << s
          << "(const float* q, long long n, float* p) { mp_poses_kernel<" << R << ">(q, n, p); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_check_" << s
          << "(const float* q, long long n, SceneDev sc, unsigned char* v, float* c, unsigned long long* w) { mp_check_kernel<" << R
          << ">(q, n, sc, v, c, w); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_cost_" << s
          << "(const float* q, long long n, SceneDev sc, float* c, unsigned char* v, float* parts, unsigned long long* w) { mp_cost_kernel<" << R
          << ">(q, n, sc, c, v, parts, w); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_sample_" << s
          << "(u64 seed, u64 off, long long n, float* q) { mp_sample_kernel<" << R << ">(seed, off, n, q, MP_SAMPLER_UNIFORM); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_halton_" << s
          << "(u64 seed, u64 off, long long n, float* q) { mp_sample_kernel<" << R << ">(seed, off, n, q, MP_SAMPLER_HALTON); }\n";
        o << "extern \"C\" __global__ void __launch_bounds__(256) mp_edges_" << s
          << "(const float* a, const float* b, int m, float res, SceneDev sc, unsigned char* ok, float* len, unsigned long long* w) { mp_edges_kernel<"
          << R << ">(a, b, m, res, sc, ok, len, w); }\n";
        if (robots.size() == 1) {  // the roadmap planner runs on a single robot's module
            o << "extern \"C\" __global__ void __launch_bounds__(MP_RM_MAX_K) mp_rm_sample_" << s
              << "(RoadmapDev g, SceneDev sc) { mp_rm_sample<" << R << ">(g, sc); }\n";
            o << "extern \"C\" __global__ void __launch_bounds__(256) mp_rm_edges_" << s
              << "(RoadmapDev g, SceneDev sc, unsigned long long* w) { mp_rm_edges<" << R << ">(g, sc, w); }\n";
            o << "extern \"C\" __global__ void __launch_bounds__(64) mp_rm_commit_" << s << "(RoadmapDev g) { mp_rm_commit<"
              << R << "::DOF>(g); }\n";
        }
    }
    // planner: dispatch on the query's robot slot (wave-uniform)
    o << "__device__ __forceinline__ void mp_load_query(const QueryDev* p, QueryDev& Q) {\n"
         "    Q.scene.ns = mp_sld(&p->scene.ns); Q.scene.nb = mp_sld(&p->scene.nb); Q.scene.nc = mp_sld(&p->scene.nc);\n"
         "    Q.scene.ncap = mp_sld(&p->scene.ncap);\n"
         "    Q.scene.sph = mp_sld(&p->scene.sph); Q.scene.box = mp_sld(&p->scene.box); Q.scene.cyl = mp_sld(&p->scene.cyl);\n"
         "    Q.scene.cap = mp_sld(&p->scene.cap); Q.scene.bnd = mp_sld(&p->scene.bnd);\n"
         "    Q.scene.bbox = mp_sld(&p->scene.bbox); Q.scene.bcyl = mp_sld(&p->scene.bcyl); Q.scene.bcap = mp_sld(&p->scene.bcap);\n"
         "    Q.tree[0] = mp_sld(&p->tree[0]); Q.tree[1] = mp_sld(&p->tree[1]);\n"
         "    Q.parent[0] = mp_sld(&p->parent[0]); Q.parent[1] = mp_sld(&p->parent[1]); Q.path = mp_sld(&p->path);\n"
         "    Q.seed = mp_sld(&p->seed); Q.robot = mp_sld(&p->robot); Q.dof = mp_sld(&p->dof); Q.cap = mp_sld(&p->cap);\n"
         "    Q.max_iters = mp_sld(&p->max_iters); Q.max_path = mp_sld(&p->max_path); Q.dof_max = mp_sld(&p->dof_max);\n"
         "    Q.sampler = mp_sld(&p->sampler);\n"
         "}\n";
    // planning kernels: one team (workgroup of W waves, W = blockDim.x / 64) per query iteration
    const std::string team_setup =
        "    __shared__ u64 team_red[2 * MP_TEAM_MAX];\n"
        "    __shared__ int team_flag;\n"
        "    Team T;\n"
        "    T.w = mp_uniform((int)(threadIdx.x >> 6));\n"
        "    T.W = mp_uniform((int)(blockDim.x >> 6));\n"
        "    T.red = team_red;\n"
        "    T.flag = &team_flag;\n"
        "    if (threadIdx.x == 0) team_flag = 0;\n";
    o << "extern \"C\" __global__ void __launch_bounds__(64 * MP_TEAM_MAX) mp_plan(PlanCtl ctl) {\n"
         "    const int lane = threadIdx.x & (MP_WAVE - 1);\n"
         "    const int n_active = mp_uniform(ctl.ctl[2]);\n"
         "    const int budget = mp_uniform(ctl.ctl[3]);\n"
         "    const int per_ticket = mp_uniform(ctl.ctl[7]) > 0 ? mp_uniform(ctl.ctl[7]) : 1;\n"
         "    const int cap = mp_uniform(ctl.ctl[8]);\n"
         "    const bool pair = mp_uniform(ctl.ctl[9]) != 0;  // two iterations per pass (32 lanes each)\n"
         "    const int cu_slot = mp_cu_slot();\n"
         "    if (n_active <= 0) return;\n"
      << team_setup <<
         "    const bool lead = T.w == 0;\n"
         "    __shared__ u64 prof_lds[MP_PROF_N];\n"
         "    u64* pl = (ctl.prof && lead) ? prof_lds : nullptr;\n"
         "    if (pl && lane == 0) for (int k = 0; k < MP_PROF_N; ++k) pl[k] = 0ull;\n"
         "    __syncthreads();\n"
         "    const u64 t_wave = mp_now(pl);\n"
         "    if (ctl.ctl[6]) {\n"
         "        // exclusive: team b alone grows query active[b] (deterministic per seed)\n"
         "        const int b = (int)blockIdx.x;\n"
         "        if (b >= n_active) return;\n"
         "        const int qi = mp_uniform(ctl.active[b]);\n"
         "        QueryState* st = ctl.st + qi;\n"
         "        QueryDev Q;\n"
         "        mp_load_query(ctl.q + qi, Q);\n"
         "        for (int it = 0; budget <= 0 || it < budget; ++it) {\n"
         "            bool more = false;\n"
         "            switch (Q.robot) {\n";
    for (size_t k = 0; k < robots.size(); k++)
        o << "                case " << k << ": more = pair ? mp_plan_pair<Robot" << k << ">(ctl, Q, st, pl, T) : mp_plan_iteration<Robot"
          << k << ">(ctl, Q, st, pl, T); break;\n";
    o << "                default: break;\n            }\n            if (!more) break;\n        }\n"
         "        mp_prof_add(pl, MP_PROF_WAVE, mp_now(pl) - t_wave);\n"
         "        mp_prof_flush(ctl, pl);\n"
         "        return;\n    }\n"
         "    for (;;) {\n"
         "        u64 ts = mp_now(pl);\n"
         "        // the lead wave takes a ticket and picks the query: its CU's query when that one runs\n"
         "        // below the cap, else round-robin over the active list, skipping finished queries with\n"
         "        // a wave-wide probe (-2: the launch is done)\n"
         "        int qi = -2;\n"
         "        if (lead) {\n"
         "            u32 t = 0;\n"
         "            if (lane == 0) t = __hip_atomic_fetch_add((u32*)ctl.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "            // the running count and the CU's query are read while the ticket is in flight\n"
         "            const int running0 = __hip_atomic_load(ctl.ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "            const int home0 = __hip_atomic_load(&mp_affinity[cu_slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "            t = (u32)__shfl((int)t, 0);\n"
         "            const int running = mp_uniform(running0);\n"
         "            // the cap grows as queries finish, so the last ones share every team of the launch\n"
         "            const int fair = running > 0 ? (int)((gridDim.x + (unsigned)running - 1u) / (unsigned)running) : 1;\n"
         "            const int capq = cap > 0 ? (cap > fair ? cap : fair) : 0;\n"
         "            if (!(budget > 0 && t >= (u32)budget) && running > 0) {\n"
         "                qi = -1;\n"
         "                bool full = false;  // a running query was passed over: it had its teams\n"
         "                const int home = mp_uniform(home0);\n"
         "                if (home >= 0 && home < n_active) {\n"
         "                    const int cand = mp_uniform(ctl.active[home]);\n"
         "                    if (mp_uniform(__hip_atomic_load(&ctl.st[cand].status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == MP_PLAN_RUNNING) {\n"
         "                        int old = 0;\n"
         "                        if (capq > 0 && lane == 0) old = __hip_atomic_fetch_add(&ctl.st[cand].workers, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "                        old = __shfl(old, 0);\n"
         "                        if (capq <= 0 || old < capq) qi = cand;\n"
         "                        else if (lane == 0) __hip_atomic_fetch_add(&ctl.st[cand].workers, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "                    }\n"
         "                }\n"
         "                int taken = -1;  // active-list position of a query the probe below takes\n"
         "                const int slot = (int)(t % (u32)n_active);\n"
         "                for (int base = 0; base < n_active && qi < 0; base += MP_WAVE) {\n"
         "                    const int p = base + lane;\n"
         "                    int a = -1, k = 0;\n"
         "                    bool run = false;\n"
         "                    if (p < n_active) {\n"
         "                        k = slot + p;\n"
         "                        if (k >= n_active) k -= n_active;\n"
         "                        a = ctl.active[k];\n"
         "                        run = __hip_atomic_load(&ctl.st[a].status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == MP_PLAN_RUNNING;\n"
         "                    }\n"
         "                    u64 b = __ballot(run);\n"
         "                    while (b) {\n"
         "                        const int cand = __shfl(a, __builtin_ctzll(b));\n"
         "                        const int pos = __shfl(k, __builtin_ctzll(b));\n"
         "                        if (capq <= 0) { qi = cand; taken = pos; break; }\n"
         "                        int old = 0;\n"
         "                        if (lane == 0) old = __hip_atomic_fetch_add(&ctl.st[cand].workers, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "                        old = __shfl(old, 0);\n"
         "                        if (old < capq) { qi = cand; taken = pos; break; }\n"
         "                        if (lane == 0) __hip_atomic_fetch_add(&ctl.st[cand].workers, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "                        full = true;\n"
         "                        b &= b - 1ull;\n"
         "                    }\n"
         "                }\n"
         "                qi = mp_uniform(qi);\n"
         "                if (taken >= 0 && lane == 0)\n"
         "                    __hip_atomic_store(&mp_affinity[cu_slot], taken, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "                // every running query it saw has its teams: this team is surplus and leaves\n"
         "                if (qi < 0 && full) qi = -2;\n"

         "            }\n"
         "        }\n"
         "        // wave-uniform: the descriptor and scene loads below then use scalar loads\n"
         "        qi = mp_uniform(mp_team_bcast(qi, T));\n"
         "        if (qi == -2) break;\n"
         "        if (qi < 0) { mp_prof_since(pl, MP_PROF_SCHED, ts); continue; }\n"

         "        QueryState* st = ctl.st + qi;\n"
         "        QueryDev Q;\n"
         "        mp_load_query(ctl.q + qi, Q);\n"
         "        mp_prof_since(pl, MP_PROF_SCHED, ts);\n"
         "        for (int rep = 0; rep < per_ticket; ++rep) {\n"
         "            bool more = false;\n"
         "            switch (Q.robot) {\n";
    for (size_t k = 0; k < robots.size(); k++)
        o << "                case " << k << ": more = pair ? mp_plan_pair<Robot" << k << ">(ctl, Q, st, pl, T) : mp_plan_iteration<Robot"
          << k << ">(ctl, Q, st, pl, T); break;\n";
    o << "                default: break;\n            }\n            if (!more) break;\n        }\n"
         "        if (cap > 0 && lead && lane == 0) __hip_atomic_fetch_add(&st->workers, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
         "    }\n"
         "    mp_prof_add(pl, MP_PROF_WAVE, mp_now(pl) - t_wave);\n"
         "    mp_prof_flush(ctl, pl);\n"
         "}\n";
    o << "extern \"C\" __global__ void __launch_bounds__(64 * MP_TEAM_MAX) mp_plan_init(PlanCtl ctl) {\n"
         "    const int b = (int)blockIdx.x;\n"
         "    const int first = ctl.ctl[4], count = ctl.ctl[5];\n"
         "    if (b >= count) return;\n"
      << team_setup <<
         "    __syncthreads();\n"
         "    const int qi = mp_uniform(ctl.active[first + b]);\n"
         "    QueryDev Q;\n"
         "    mp_load_query(ctl.q + qi, Q);\n"
         "    for (int j = 0; j < 16; ++j) { Q.start[j] = ctl.q[qi].start[j]; Q.goal[j] = ctl.q[qi].goal[j]; }\n"
         "    switch (Q.robot) {\n";
    for (size_t k = 0; k < robots.size(); k++)
        o << "        case " << k << ": mp_plan_init_one<Robot" << k << ">(ctl, Q, ctl.st + qi, T); break;\n";
    o << "        default: break;\n    }\n}\n";
    return o.str();
}

}  // namespace mp
