// Pointwise physics of the pseudo-time step, written once for the host and
// the gfx950 device (HD = __host__ __device__).  Templates take the
// arithmetic type R of the face/flux work (double on the FP64 path, float on
// the mixed path); the IF97 saturation line, the LU-SGS diagonal inverse and
// every norm stay in double.
//
// Reference formulas (kwonhwi1/3dcode):
//   IF97 region 1/2/metastable-2    src/eos_module.f90:1331-1461
//   IAPS-85 transport, ideal gas    src/eos_module.f90:1236-1282, 1463-1659
//   saturation line                 src/eos_module.f90:2266-2279
//   mixture rules                   src/eos_module.f90:1064-1118, 1176-1234
//   MUSCL-TVD / MLP / limiters      src/muscl_module.f90:104-413
//   preconditioned Roe, AUSM+-up    src/flux_module.f90:198-276, 548-695
//   viscous flux + SST F1           src/vsflux_module.f90:182-425
//   Merkle source                   src/cav_module.f90:136-171
//   SST source / eddy viscosity     src/turbsource_module.f90:269-414,
//                                   src/eddy_module.f90:135-187
//   split Jacobian (applied as a    src/jacobian_module.f90:182-465
//     structured mat-vec)
//   diagonal inverse getx           src/lhs_module.f90:346-534
//   local time step                 src/timestep_module.f90:156-198
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define HD __host__ __device__ __forceinline__

namespace cav {

constexpr int NPV = 9;

// compact cell properties kept per cell (the dv/tv entries the path reads)
enum { D_RHO = 0, D_H, D_C2, D_RP, D_RT, D_RY1, D_RY2, D_HP, D_HT, D_HY1, D_HY2, D_MU, D_KAP, ND };

struct Par {
    int npv, ntv, iturb, nscheme, nmuscl, nlim, local, prec, ncav, nsteady, cav_clip, ides;
    int uref_seed;   // dual lane + 1 that carries d/d(uref) in the adjoint's passes (0: none)
    int dpref_seed, tref_seed;   // the same for the mass-transfer scales dp_ref and t_ref
    double cfl, pref, uref, str, dt_phy, c_v, c_c, dp_ref, t_ref, t_crit;
    double bc_pv[9], bc_pressure, emutref, bdf_b;
};

template <class R> HD R mx(R a, R b) { return (b > a) ? b : a; }
template <class R> HD R mn(R a, R b) { return (b < a) ? b : a; }

// ------------------------------------------------------------ dual numbers
// (the double/float overloads stay visible beside the dual ones below)
using ::sqrt;
using ::fabs;
using ::exp;
using ::log;
using ::tanh;
using ::pow;
using ::signbit;

// Forward-mode dual number with W tangent lanes.  The adjoint (VJP) kernels
// instantiate the templates below with it: each pass seeds up to W inputs of
// a pointwise function and reads the W directional derivatives of every
// output, so the derivative follows the same branches, clamps and
// min/max selections as the value.
template <int W> struct Dn {
    double v, d[W];
    HD Dn() {}
    HD Dn(double x) : v(x)
    {
#pragma unroll
        for (int q = 0; q < W; q++) d[q] = 0.0;
    }
    HD explicit operator double() const { return v; }
    HD static Dn seed(double x, int lane)
    {
        Dn r(x);
        if (lane >= 0 && lane < W) r.d[lane] = 1.0;
        return r;
    }
};
template <int W> HD Dn<W> dn_map(const Dn<W> &a, double v, double da)
{
    Dn<W> r;
    r.v = v;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = da * a.d[q];
    return r;
}
template <int W> HD Dn<W> operator-(const Dn<W> &a) { return dn_map(a, -a.v, -1.0); }
template <int W> HD Dn<W> operator+(const Dn<W> &a, const Dn<W> &b)
{
    Dn<W> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = a.d[q] + b.d[q];
    return r;
}
template <int W> HD Dn<W> operator-(const Dn<W> &a, const Dn<W> &b)
{
    Dn<W> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = a.d[q] - b.d[q];
    return r;
}
template <int W> HD Dn<W> operator*(const Dn<W> &a, const Dn<W> &b)
{
    Dn<W> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
template <int W> HD Dn<W> operator/(const Dn<W> &a, const Dn<W> &b)
{
    Dn<W> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = (a.d[q] - r.v * b.d[q]) * ib;
    return r;
}
template <int W> HD Dn<W> operator+(const Dn<W> &a, double b) { return a + Dn<W>(b); }
template <int W> HD Dn<W> operator+(double a, const Dn<W> &b) { return Dn<W>(a) + b; }
template <int W> HD Dn<W> operator-(const Dn<W> &a, double b) { return a - Dn<W>(b); }
template <int W> HD Dn<W> operator-(double a, const Dn<W> &b) { return Dn<W>(a) - b; }
template <int W> HD Dn<W> operator*(const Dn<W> &a, double b) { return dn_map(a, a.v * b, b); }
template <int W> HD Dn<W> operator*(double a, const Dn<W> &b) { return dn_map(b, a * b.v, a); }
template <int W> HD Dn<W> operator/(const Dn<W> &a, double b) { return a / Dn<W>(b); }
template <int W> HD Dn<W> operator/(double a, const Dn<W> &b) { return Dn<W>(a) / b; }
#define CAV_DN_CMP(OP)                                                                        \
    template <int W> HD bool operator OP(const Dn<W> &a, const Dn<W> &b) { return a.v OP b.v; } \
    template <int W> HD bool operator OP(const Dn<W> &a, double b) { return a.v OP b; }         \
    template <int W> HD bool operator OP(double a, const Dn<W> &b) { return a OP b.v; }
CAV_DN_CMP(<)
CAV_DN_CMP(>)
CAV_DN_CMP(<=)
CAV_DN_CMP(>=)
CAV_DN_CMP(==)
CAV_DN_CMP(!=)
#undef CAV_DN_CMP
// at an argument of exactly 0 the derivatives below are taken as 0 (the
// rate and blending terms reach 0 only through a max(., 0) or a zero mass
// fraction, where the one-sided derivative that matters is 0)
template <int W> HD Dn<W> sqrt(const Dn<W> &a)
{
    double s = ::sqrt(a.v);
    return dn_map(a, s, s > 0.0 ? 0.5 / s : 0.0);
}
template <int W> HD Dn<W> fabs(const Dn<W> &a) { return dn_map(a, ::fabs(a.v), a.v < 0.0 ? -1.0 : 1.0); }
template <int W> HD Dn<W> exp(const Dn<W> &a) { double e = ::exp(a.v); return dn_map(a, e, e); }
template <int W> HD Dn<W> log(const Dn<W> &a) { return dn_map(a, ::log(a.v), 1.0 / a.v); }
template <int W> HD Dn<W> tanh(const Dn<W> &a) { double t = ::tanh(a.v); return dn_map(a, t, 1.0 - t * t); }
template <int W> HD Dn<W> pow(const Dn<W> &a, double b)
{
    double p = ::pow(a.v, b);
    return dn_map(a, p, b == 0.0 ? 0.0 : (a.v == 0.0 ? (b == 1.0 ? 1.0 : 0.0) : b * ::pow(a.v, b - 1.0)));
}
template <int W> HD Dn<W> pow(const Dn<W> &a, const Dn<W> &b)
{
    Dn<W> r;
    r.v = ::pow(a.v, b.v);
    const double la = a.v > 0.0 ? ::log(a.v) : 0.0, ea = b.v * ::pow(a.v, b.v - 1.0);
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = ea * a.d[q] + r.v * la * b.d[q];
    return r;
}
template <int W> HD bool signbit(const Dn<W> &a) { return ::signbit(a.v); }

template <class R> struct IsDual { static constexpr bool v = false; };
template <int W> struct IsDual<Dn<W>> { static constexpr bool v = true; };

// min(a, b) of a mass-transfer rate a and its clip b, both proportional to a
// mass fraction s >= 0 (y1, or 1 - y1 - y2).  The value is mn's.  At a tie
// (typically s = 0, where both vanish) a dual takes, per lane, the slope of
// the smaller function on the feasible side: min of the two slopes where the
// lane moves s up, max where it moves s down.
template <class R> HD R mn_rate(R a, R b, R) { return mn(a, b); }
template <int W> HD Dn<W> mn_rate(const Dn<W> &a, const Dn<W> &b, const Dn<W> &s)
{
    if (a.v != b.v) return mn(a, b);
    Dn<W> r = a;
#pragma unroll
    for (int q = 0; q < W; q++) r.d[q] = s.d[q] >= 0.0 ? ::fmin(a.d[q], b.d[q]) : ::fmax(a.d[q], b.d[q]);
    return r;
}

// the preconditioning reference speed as R; a dual carries d/d(uref) on the
// lane Par::uref_seed names
template <class R> struct RefSpeed {
    HD static R of(const Par &P) { return R(P.uref); }
};
template <int W> struct RefSpeed<Dn<W>> {
    HD static Dn<W> of(const Par &P) { return Dn<W>::seed(P.uref, P.uref_seed - 1); }
};
// the mass-transfer scales dp_ref = 2/(rho_ref uref^2) and t_ref = uref/l
// (src/cav_module.f90:81-82) as R, seeded like RefSpeed
template <class R> struct CavScale {
    HD static R dp(const Par &P) { return R(P.dp_ref); }
    HD static R t(const Par &P) { return R(P.t_ref); }
};
template <int W> struct CavScale<Dn<W>> {
    HD static Dn<W> dp(const Par &P) { return Dn<W>::seed(P.dp_ref, P.dpref_seed - 1); }
    HD static Dn<W> t(const Par &P) { return Dn<W>::seed(P.t_ref, P.tref_seed - 1); }
};

// the face EOS skips a phase whose mass fraction is zero (its terms are exact
// zeros); a dual keeps the phase when the fraction's derivative is non-zero
template <class R> HD bool has_mass(R y) { return y != R(0); }
template <int W> HD bool has_mass(const Dn<W> &y)
{
    bool nz = y.v != 0.0;
#pragma unroll
    for (int q = 0; q < W; q++) nz = nz || y.d[q] != 0.0;
    return nz;
}


// ------------------------------------------------------------ IF97 tables
// IAPWS-IF97 (2007) region 1 (table 2), region 2 ideal/residual (tables 10,
// 11) and metastable-vapour residual (table 16) coefficients.
struct IJN { int i, j; double n; };
// Coefficient tables live in device memory in the device pass and as plain
// constants in the host pass (the host never reads a device variable's shadow).
#if defined(__HIP_DEVICE_COMPILE__)
#define CAV_TABLE __device__ constexpr
#else
#define CAV_TABLE constexpr
#endif
CAV_TABLE IJN R1C[34] = {
    {0, -2, 0.14632971213167e0}, {0, -1, -0.84548187169114e0}, {0, 0, -0.37563603672040e1},
    {0, 1, 0.33855169168385e1}, {0, 2, -0.95791963387872e0}, {0, 3, 0.15772038513228e0},
    {0, 4, -0.16616417199501e-1}, {0, 5, 0.81214629983568e-3}, {1, -9, 0.28319080123804e-3},
    {1, -7, -0.60706301565874e-3}, {1, -1, -0.18990068218419e-1}, {1, 0, -0.32529748770505e-1},
    {1, 1, -0.21841717175414e-1}, {1, 3, -0.52838357969930e-4}, {2, -3, -0.47184321073267e-3},
    {2, 0, -0.30001780793026e-3}, {2, 1, 0.47661393906987e-4}, {2, 3, -0.44141845330846e-5},
    {2, 17, -0.72694996297594e-15}, {3, -4, -0.31679644845054e-4}, {3, 0, -0.28270797985312e-5},
    {3, 6, -0.85205128120103e-9}, {4, -5, -0.22425281908000e-5}, {4, -2, -0.65171222895601e-6},
    {4, 10, -0.14341729937924e-12}, {5, -8, -0.40516996860117e-6}, {8, -11, -0.12734301741641e-8},
    {8, -6, -0.17424871230634e-9}, {21, -29, -0.68762131295531e-18}, {23, -31, 0.14478307828521e-19},
    {29, -38, 0.26335781662795e-22}, {30, -39, -0.11947622640071e-22}, {31, -40, 0.18228094581404e-23},
    {32, -41, -0.93537087292458e-25}};
CAV_TABLE int R2J0[9] = {0, 1, -5, -4, -3, -2, -1, 2, 3};
CAV_TABLE double R2N0[9] = {-0.96927686500217e1, 0.10086655968018e2, -0.56087911283020e-2,
    0.71452738081455e-1, -0.40710498223928e0, 0.14240819171444e1, -0.43839511319450e1, -0.28408632460772e0,
    0.21268463753307e-1};
CAV_TABLE double R2MN0[2] = {-0.96937268393049e1, 0.10087275970006e2};
CAV_TABLE IJN R2C[43] = {
    {1, 0, -0.17731742473213e-2}, {1, 1, -0.17834862292358e-1}, {1, 2, -0.45996013696365e-1},
    {1, 3, -0.57581259083432e-1}, {1, 6, -0.50325278727930e-1}, {2, 1, -0.33032641670203e-4},
    {2, 2, -0.18948987516315e-3}, {2, 4, -0.39392777243355e-2}, {2, 7, -0.43797295650573e-1},
    {2, 36, -0.26674547914087e-4}, {3, 0, 0.20481737692309e-7}, {3, 1, 0.43870667284435e-6},
    {3, 3, -0.32277677238570e-4}, {3, 6, -0.15033924542148e-2}, {3, 35, -0.40668253562649e-1},
    {4, 1, -0.78847309559367e-9}, {4, 2, 0.12790717852285e-7}, {4, 3, 0.48225372718507e-6},
    {5, 7, 0.22922076337661e-5}, {6, 3, -0.16714766451061e-10}, {6, 16, -0.21171472321355e-2},
    {6, 35, -0.23895741934104e2}, {7, 0, -0.59059564324270e-17}, {7, 11, -0.12621808899101e-5},
    {7, 25, -0.38946842435739e-1}, {8, 8, 0.11256211360459e-10}, {8, 36, -0.82311340897998e1},
    {9, 13, 0.19809712802088e-7}, {10, 4, 0.10406965210174e-18}, {10, 10, -0.10234747095929e-12},
    {10, 14, -0.10018179379511e-8}, {16, 29, -0.80882908646985e-10}, {16, 50, 0.10693031879409e0},
    {18, 57, -0.33662250574171e0}, {20, 20, 0.89185845355421e-24}, {20, 35, 0.30629316876232e-12},
    {20, 48, -0.42002467698208e-5}, {21, 21, -0.59056029685639e-25}, {22, 53, 0.37826947613457e-5},
    {23, 39, -0.12768608934681e-14}, {24, 26, 0.73087610595061e-28}, {24, 40, 0.55414715350778e-16},
    {24, 58, -0.94369707241210e-6}};
CAV_TABLE IJN R2MC[13] = {
    {1, 0, -0.73362260186506e-2}, {1, 2, -0.88223831943146e-1}, {1, 5, -0.72334555213245e-1},
    {1, 11, -0.40813178534455e-2}, {2, 1, 0.20097803380207e-2}, {2, 7, -0.53045921898642e-1},
    {2, 16, -0.76190409086970e-2}, {3, 4, -0.63498037657313e-2}, {3, 16, -0.86043093028588e-1},
    {4, 7, 0.75321581522770e-2}, {4, 10, -0.79238375446139e-2}, {5, 9, -0.22888160778447e-3},
    {5, 10, -0.26456501482810e-2}};
// IAPS-85 viscosity / conductivity coefficients
CAV_TABLE double TPH[4] = {1.0, 0.978197, 0.579829, -0.202354};
CAV_TABLE double TPL[4] = {1.0, 6.978267, 2.599096, -0.998254};
CAV_TABLE double TPHH[6][7] = {
    {0.5132047, 0.2151778, -0.2818107, 0.1778064, -0.0417661, 0.0, 0.0},
    {0.3205656, 0.7317883, -1.070786, 0.460504, 0.0, -0.01578386, 0.0},
    {0.0, 1.241044, -1.263184, 0.2340379, 0.0, 0.0, 0.0},
    {0.0, 1.476783, 0.0, -0.4924179, 0.1600435, 0.0, -0.003629481},
    {-0.7782567, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.1885447, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
CAV_TABLE double TPLL[5][6] = {
    {1.3293046, -0.40452437, 0.24409490, 0.018660751, -0.12961068, 0.044809953},
    {1.7018363, -2.2156845, 1.6511057, -0.76736002, 0.37283344, -0.11203160},
    {5.2246158, -10.124111, 4.9874687, -0.27297694, -0.43083393, 0.13333849},
    {8.7127675, -9.5000611, 4.3786606, -0.91783782, 0.0, 0.0},
    {-1.8525999, 0.93404690, 0.0, 0.0, 0.0, 0.0}};

// Folded constants.  Five of them are default-REAL literals in the reference
// (src/eos_module.f90:1367,1370,1371,1457) and are used at float precision.
constexpr double C_RHO1 = (double)25.841246054191803727f;   // p*/(R T*) region 1
constexpr double C_RT1 = (double)0.01864447767257706f;
constexpr double C_HP1 = (double)0.03869782431941923775f;
constexpr double C_RT2A = (double)0.00185185185185185185f;  // 1/540
constexpr double C_RT2B = (double)0.0074304703986495518f;

template <class R> struct E2 { R rho, h, drdp, drdt, dhdp, dhdt; };

// IF97 region-4 saturation pressure (src/eos_module.f90:2266-2279): double
// on every evaluation path (the face EOS rounds its temperature to double for
// it); a dual only in the adjoint's cavitation source
template <class R> HD R sat_p(R t)
{
    R tt = mn(mx(t, R(273.15)), R(647.096));
    R th = tt - 0.23855557567849e0 / (tt - 0.65017534844798e3);
    R a = th * th + 0.11670521452767e4 * th - 0.72421316703206e6;
    R b = -0.17073846940092e2 * (th * th) + 0.12020824702470e5 * th - 0.32325550322333e7;
    R c = 0.14915108613530e2 * (th * th) - 0.48232657361591e4 * th + 0.40511340542057e6;
    R x = 2.0 * c / (-b + sqrt(b * b - 4.0 * a * c));
    R x2 = x * x;
    return 1.0e6 * (x2 * x2);
}

// IF97 region-4 saturation temperature (src/eos_module.f90:2340-2354)
template <class R> HD R sat_t(R p)
{
    R beta = pow(p / 1.0e6, 0.25);
    R e = beta * beta - 0.17073846940092e2 * beta + 0.14915108613530e2;
    R f = 0.11670521452767e4 * (beta * beta) + 0.12020824702470e5 * beta - 0.48232657361591e4;
    R g = -0.72421316703206e6 * (beta * beta) - 0.32325550322333e7 * beta + 0.40511340542057e6;
    R d = 2.0 * g / (-f - sqrt(f * f - 4.0 * e * g));
    R s = 0.65017534844798e3 + d;
    return 0.5 * (0.65017534844798e3 + d - sqrt(s * s - 4.0 * (-0.23855557567849e0 + 0.65017534844798e3 * d)));
}

// surface tension of water against its vapour (src/eos_module.f90:2430-2438);
// constant integer powers as gfortran expands them: x**3 = (x*x)*x,
// x**4 = (x*x)*(x*x), x**5 = ((x*x)*x)*(x*x)
template <class R> HD R h2o_sigma(R t)
{
    R t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t3 * t2;
    return 0.0881130429998226 + 0.0000616475914044782 * t - 4.4860167748009e-7 * t2 + 2.53682468718936e-10 * t3
           - 2.62005668908913e-13 * t4 + 2.97878955533757e-16 * t5;
}

// real**integer as gfortran evaluates it with a run-time exponent (libgcc
// __powidf2: binary exponentiation, reciprocal for negative powers); the
// exponents here are compile-time table entries, so the loop unrolls into
// the same multiplication sequence.
template <class R> HD R powi(R x, int m)
{
    unsigned n = m < 0 ? 0u - (unsigned)m : (unsigned)m;
    R y = (n % 2u) ? x : R(1);
    while (n >>= 1) {
        x = x * x;
        if (n % 2u) y = y * x;
    }
    return m < 0 ? R(1) / y : y;
}

// Region 1 (src/eos_module.f90:1331-1374), term by term in the reference's
// operation order so the FP64 path reproduces its rounding.
template <class R> HD E2<R> liquid(R p, R t)
{
    constexpr int EOS_UNROLL = IsDual<R>::v ? 1 : 64;   // dual instantiations stay rolled
    R pi = p * R(0.00000006049607);
    R tau = R(1386.0) / t;
    R pi1 = R(1) / (R(7.1) - pi);
    R tau1 = R(1) / (tau - R(1.222));
    R gt = 0, gp = 0, gpp = 0, gtt = 0, gpt = 0;
#pragma unroll EOS_UNROLL
    for (int k = 0; k < 34; k++) {
        const int I = R1C[k].i, J = R1C[k].j;
        R g = R(R1C[k].n) * powi(R(7.1) - pi, I) * powi(tau - R(1.222), J);
        R g1 = R(J) * g * tau1;
        R g2 = -(R(I) * g * pi1);
        R g3 = -(R(I - 1) * g2 * pi1);
        R g4 = R(J - 1) * g1 * tau1;
        R g5 = R(J) * g2 * tau1;
        gt = gt + g1;
        gp = gp + g2;
        gpp = gpp + g3;
        gtt = gtt + g4;
        gpt = gpt + g5;
    }
    R gp1 = R(1) / gp;
    R tau3 = tau * tau * tau;
    E2<R> e;
    e.rho = R(C_RHO1) * tau * gp1;
    e.h = R(639675.036) * gt;
    e.drdp = -(gpp * R(1.5632937721834122e-6) * tau * (gp1 * gp1));
    e.drdt = -(e.rho * R(7.215007215007215e-4) * tau) + R(C_RT1) * gpt * tau3 * (gp1 * gp1);
    e.dhdp = R(C_HP1) * gpt;
    e.dhdt = -(R(461.526) * (tau * tau) * gtt);
    return e;
}

// Region 2 (src/eos_module.f90:1376-1461) with the metastable-vapour
// equation whenever p > p_sat(T), in the reference's operation order.
template <class R> HD E2<R> vapour(R p, R t, double psat)
{
    constexpr int EOS_UNROLL = IsDual<R>::v ? 1 : 64;   // dual instantiations stay rolled
    R pi = p * R(1.0e-6);
    R tau = R(540.0) / t;
    R pi1 = R(1.0e6) / p;
    R tau1 = R(1) / tau;
    R tau2 = R(1) / (tau - R(0.5));
    R g0t = 0, g0tt = 0;
    R grt = 0, grp = 0, grpp = 0, grtt = 0, grpt = 0;
    const bool meta = (double)p > psat;
#pragma unroll EOS_UNROLL
    for (int k = 0; k < 9; k++) {
        const int J = R2J0[k];
        double n = R2N0[k];
        if (meta && k < 2) n = R2MN0[k];
        R g0 = R(n) * powi(tau, J);
        R g01 = R(J) * g0 * tau1;
        R g02 = R(J - 1) * g01 * tau1;
        g0t = g0t + g01;
        g0tt = g0tt + g02;
    }
    auto term = [&](const IJN c) {
        R gr = R(c.n) * powi(pi, c.i) * powi(tau - R(0.5), c.j);
        R gr1 = R(c.j) * gr * tau2;
        R gr2 = R(c.i) * gr * pi1;
        R gr3 = R(c.i - 1) * gr2 * pi1;
        R gr4 = R(c.j - 1) * gr1 * tau2;
        R gr5 = R(c.j) * gr2 * tau2;
        grt = grt + gr1;
        grp = grp + gr2;
        grpp = grpp + gr3;
        grtt = grtt + gr4;
        grpt = grpt + gr5;
    };
    if (meta) {
#pragma unroll EOS_UNROLL
        for (int k = 0; k < 13; k++) term(R2MC[k]);
    } else {
#pragma unroll EOS_UNROLL
        for (int k = 0; k < 43; k++) term(R2C[k]);
    }
    R gp1 = R(1) / (pi1 + grp);
    R tau3 = tau * tau * tau;
    E2<R> e;
    e.rho = R(4.01245401527075799) * tau * gp1;
    e.h = R(249224.04) * (g0t + grt);
    e.drdp = -((-(pi1 * pi1) + grpp) * R(4.01245401527075799e-6) * tau * (gp1 * gp1));
    e.drdt = -(e.rho * R(C_RT2A) * tau) + R(C_RT2B) * (R(0) + grpt) * tau3 * (gp1 * gp1);
    e.dhdp = R(0.24922404) * (R(0) + grpt);
    e.dhdt = -(R(461.526) * (tau * tau) * (g0tt + grtt));
    return e;
}

template <class R> HD E2<R> gas(R p, R t)
{
    const R r1 = R(0.00348383500557);
    R tau = R(1) / t;
    E2<R> e;
    e.rho = p * r1 * tau;
    e.h = R(1004.64) * t;
    e.drdp = r1 * tau;
    e.drdt = -(p * r1 * (tau * tau));
    e.dhdp = R(0);
    e.dhdt = R(1004.64);
    return e;
}

// IAPS-85 mu, kappa of a water phase, no critical enhancement
// (src/eos_module.f90:1508-1537), in the reference's operation order.
template <class R> HD void iaps85(R t, R rho, R &mu, R &kap)
{
    constexpr int EOS_UNROLL = IsDual<R>::v ? 1 : 64;   // dual instantiations stay rolled
    R tt = t * R(0.00154495032985);
    R rr = mx(R(1.0e-3), mn(R(1332.4), rho)) * R(0.00314699949333);
    R tt1 = R(647.27) / t;
    R mu0 = sqrt(tt) / (R(TPH[0]) + R(TPH[1]) * tt1 + R(TPH[2]) * (tt1 * tt1) + R(TPH[3]) * (tt1 * tt1 * tt1));
    R mu1 = 0;
#pragma unroll EOS_UNROLL
    for (int i = 0; i <= 5; i++)
#pragma unroll EOS_UNROLL
        for (int j = 0; j <= 6; j++)
            mu1 = mu1 + R(TPHH[i][j]) * powi(tt1 - R(1), i) * powi(rr - R(1), j);
    mu1 = exp(rr * mu1);
    mu = R(55.071e-6) * mu0 * mu1 * R(1);
    R lam0 = sqrt(tt) / (R(TPL[0]) + R(TPL[1]) * tt1 + R(TPL[2]) * (tt1 * tt1) + R(TPL[3]) * (tt1 * tt1 * tt1));
    R lam1 = 0;
#pragma unroll EOS_UNROLL
    for (int i = 0; i <= 4; i++)
#pragma unroll EOS_UNROLL
        for (int j = 0; j <= 5; j++)
            lam1 = lam1 + R(TPLL[i][j]) * powi(tt1 - R(1), i) * powi(rr - R(1), j);
    lam1 = exp(rr * lam1);
    kap = R(0.4945) * (lam0 * lam1 + R(0));
}

// Mixture of liquid, vapour and ideal gas (fluid = IF97, ngas = ideal).
// d[ND] receives the compact properties; full (if non-null) the 18 dv entries.
// prop = compute mu, kappa (deteos_prop) as d[D_MU], d[D_KAP].
template <class R> HD void mixture_eos(R p, R t, R y1, R y2, R *d, bool prop, R *full = nullptr)
{
    double psat = sat_p((double)t);
    E2<R> l = liquid(p, t), v = vapour(p, t, psat), g = gas(p, t);
    R il = R(1) / l.rho, iv = R(1) / v.rho, ig = R(1) / g.rho;
    R yl = R(1) - y1 - y2;
    R rho = R(1) / (yl * il + y1 * iv + y2 * ig);
    R r2 = rho * rho;
    d[D_RHO] = rho;
    d[D_H] = yl * l.h + y1 * v.h + y2 * g.h;
    d[D_RP] = r2 * (yl * (il * il) * l.drdp + y1 * (iv * iv) * v.drdp + y2 * (ig * ig) * g.drdp);
    d[D_RT] = r2 * (yl * (il * il) * l.drdt + y1 * (iv * iv) * v.drdt + y2 * (ig * ig) * g.drdt);
    d[D_RY1] = r2 * (il - iv);
    d[D_RY2] = r2 * (il - ig);
    d[D_HP] = yl * l.dhdp + y1 * v.dhdp + y2 * g.dhdp;
    d[D_HT] = yl * l.dhdt + y1 * v.dhdt + y2 * g.dhdt;
    d[D_HY1] = v.h - l.h;
    d[D_HY2] = g.h - l.h;
    R c2 = rho * d[D_HT] / (rho * d[D_RP] * d[D_HT] + d[D_RT] * (R(1) - rho * d[D_HP]));
    if (c2 <= R(0)) {   // src/eos_module.f90:1104-1111
        R cl = (l.rho * l.drdp * l.dhdt + l.drdt * (R(1) - l.rho * l.dhdp)) / (l.rho * l.dhdt);
        R cv = (v.rho * v.drdp * v.dhdt + v.drdt * (R(1) - v.rho * v.dhdp)) / (v.rho * v.dhdt);
        R cg = (g.rho * g.drdp * g.dhdt + g.drdt * (R(1) - g.rho * g.dhdp)) / (g.rho * g.dhdt);
        c2 = R(1) / (r2 * (yl * cl * (il * il) + y1 * cv * (iv * iv) + y2 * cg * (ig * ig)));
    }
    d[D_C2] = c2;
    if (prop) {
        R ml, kl, mv, kv;
        iaps85(t, l.rho, ml, kl);
        iaps85(t, v.rho, mv, kv);
        R mg = R(1.716e-5) * pow(t / R(273.11), R(1.5)) * R(383.67) / (t + R(110.56));
        R kg = R(1395.333333) * mg;
        d[D_MU] = rho * (ml * yl * il + mv * y1 * iv + mg * y2 * ig);
        d[D_KAP] = rho * (kl * yl * il + kv * y1 * iv + kg * y2 * ig);
    }
    if (full) {
        full[0] = rho; full[1] = d[D_H]; full[2] = l.rho; full[3] = v.rho; full[4] = g.rho; full[5] = c2;
        full[6] = d[D_RP]; full[7] = d[D_RT]; full[8] = d[D_RY1]; full[9] = d[D_RY2];
        full[10] = d[D_HP]; full[11] = d[D_HT]; full[12] = d[D_HY1]; full[13] = d[D_HY2];
        full[14] = v.drdp; full[15] = v.drdt; full[16] = l.drdp; full[17] = l.drdt;
    }
}

// Density and static enthalpy of a face state (all the flux reads from the
// reconstructed states).  Phases with zero mass fraction contribute exact
// zeros in the reference's mixture sums, so they are skipped.
template <class R> HD void face_rho_h(R p, R t, R y1, R y2, R &rho, R &h)
{
    E2<R> l = liquid(p, t);
    R yl = R(1) - y1 - y2;
    R iv = 0, ig = 0, hv = 0, hg = 0;
    if (has_mass(y1)) { E2<R> v = vapour(p, t, sat_p((double)t)); iv = R(1) / v.rho; hv = v.h; }
    if (has_mass(y2)) { E2<R> g = gas(p, t); ig = R(1) / g.rho; hg = g.h; }
    rho = R(1) / (yl * (R(1) / l.rho) + y1 * iv + y2 * ig);
    h = yl * l.h + y1 * hv + y2 * hg;
}

// Mixture c^2 of the Roe-averaged state (src/flux_module.f90:237,588).
template <class R> HD R face_c2(R p, R t, R y1, R y2)
{
    E2<R> l = liquid(p, t), v{}, g{};
    const bool hv = has_mass(y1), hg = has_mass(y2);
    if (hv) v = vapour(p, t, sat_p((double)t));
    if (hg) g = gas(p, t);
    R il = R(1) / l.rho, iv = hv ? R(1) / v.rho : R(0), ig = hg ? R(1) / g.rho : R(0);
    R yl = R(1) - y1 - y2;
    R rho = R(1) / (yl * il + y1 * iv + y2 * ig);
    R r2 = rho * rho;
    R rp = r2 * (yl * (il * il) * l.drdp + (hv ? y1 * (iv * iv) * v.drdp : R(0)) + (hg ? y2 * (ig * ig) * g.drdp : R(0)));
    R rt = r2 * (yl * (il * il) * l.drdt + (hv ? y1 * (iv * iv) * v.drdt : R(0)) + (hg ? y2 * (ig * ig) * g.drdt : R(0)));
    R hp = yl * l.dhdp + (hv ? y1 * v.dhdp : R(0)) + (hg ? y2 * g.dhdp : R(0));
    R ht = yl * l.dhdt + (hv ? y1 * v.dhdt : R(0)) + (hg ? y2 * g.dhdt : R(0));
    R c2 = rho * ht / (rho * rp * ht + rt * (R(1) - rho * hp));
    if (c2 <= R(0)) {
        R cl = (l.rho * l.drdp * l.dhdt + l.drdt * (R(1) - l.rho * l.dhdp)) / (l.rho * l.dhdt);
        R cv = hv ? (v.rho * v.drdp * v.dhdt + v.drdt * (R(1) - v.rho * v.dhdp)) / (v.rho * v.dhdt) : R(0);
        R cg = hg ? (g.rho * g.drdp * g.dhdt + g.drdt * (R(1) - g.rho * g.dhdp)) / (g.rho * g.dhdt) : R(0);
        c2 = R(1) / (r2 * (yl * cl * (il * il) + y1 * cv * (iv * iv) + y2 * cg * (ig * ig)));
    }
    return c2;
}

// ------------------------------------------------------------ preconditioning
template <class R> HD R sndp2(const Par &P, R snd2, R u2)
{
    const R ur = RefSpeed<R>::of(P);
    R ur2 = ur * ur;
    if (P.prec == 0) return snd2;
    if (P.prec == 1) return mn(snd2, mx(u2, ur2));
    return mn(snd2, mx(mx(u2, ur2), R(P.str) * R(P.str)));
}

// the dissipation's speed of sound: steady preconditioning whenever any is on
// (getsndp2_du, src/flux_module.f90:107-116)
template <class R> HD R sndp2_du(const Par &P, R snd2, R u2)
{
    if (P.prec == 0) return snd2;
    const R ur = RefSpeed<R>::of(P);
    return mn(snd2, mx(u2, ur * ur));
}

// ------------------------------------------------------------ reconstruction
// x[ll][n]: 38-point stencil (ll = Fortran index - 1)
template <class R> struct LimState { R r[2], r1[2], r2[2], alp[2]; };

template <class R> HD R limiter(int nlim, const LimState<R> &m, int s)
{
    R r = m.r[s], a = m.alp[s], c;
    switch (nlim) {
    case 0: return R(1);
    case 1: return mx(R(0), mn(R(1), r));
    case 2: return mx(mx(R(0), mn(R(1), a * r)), mn(a, r));
    case 3: c = (r + fabs(r)) / (R(1) + fabs(r)); return mx(R(0), mn(mn(a, a * r), c));
    case 4: c = (R(1) + r) / R(2); return mx(R(0), mn(mn(a, a * r), c));
    case 5: c = (r * r + r) / (R(1) + r * r); return mx(mx(R(0), mn(c * r, R(1))), mn(r, c));
    case 6: c = mn(R(1.5), a); return mx(mx(R(0), mn(c * r, R(1))), mn(r, c));
    case 7: c = (R(1) + R(2) * r) / R(3); return mx(R(0), mn(mn(a, a * r), c));
    case 8: c = (-(R(2) * m.r1[s]) + R(11) + R(24) * r - R(3) * m.r2[s]) / R(30); return mx(R(0), mn(mn(a, a * r), c));
    case 9: return (R(1) + R(2) * r) / R(3);
    case 10: return (-(R(2) * m.r1[s]) + R(11) + R(24) * r - R(3) * m.r2[s]) / R(30);
    }
    return R(1);
}

// One variable of the MUSCL-TVD (mlp=false) or MLP reconstruction.
// xs(ll) gives the stencil value of this variable at Fortran index ll.
template <class R, class F> HD void muscl_var(int nlim, bool mlp, F xs, R &xl, R &xr)
{
    const R eps = R(1.0e-30);
    R x9 = xs(9), x10 = xs(10), x23 = xs(23), x32 = xs(32);
    R dq = x10 - x9, dqm = x9 - x23, dqm1 = x23 - xs(37), dqp = x32 - x10, dqp1 = xs(38) - x32;
    LimState<R> m;
    R dqmm = R(1) / (dqm + eps);
    m.r[0] = dq * dqmm; m.r1[0] = dqm1 * dqmm; m.r2[0] = dqp * dqmm;
    R dqpp = R(1) / (dqp + eps);
    m.r[1] = dq * dqpp; m.r1[1] = dqp1 * dqpp; m.r2[1] = dqm * dqpp;
    if (!mlp) {
        m.alp[0] = R(2); m.alp[1] = R(2);
    } else {
        const R eps2 = R(1.0e-3);
        R rxy_l = fabs((xs(11) - xs(7)) / (x10 - x23 + eps));
        R rxz_l = fabs((xs(15) - xs(3)) / (x10 - x23 + eps));
        R rxy_r = fabs((xs(12) - xs(8)) / (x32 - x9 + eps));
        R rxz_r = fabs((xs(16) - xs(4)) / (x32 - x9 + eps));
        // 18-point min/max over the two cells of the face and their rings
        R a18 = xs(1), b18 = xs(1);
#pragma unroll
        for (int q = 2; q <= 18; q++) { a18 = mn(a18, xs(q)); b18 = mx(b18, xs(q)); }
        // the nine upwind-side planes ll = 19..27 and 28..36
        R aL = xs(19), bL = xs(19), aR = xs(28), bR = xs(28);
#pragma unroll
        for (int q = 20; q <= 27; q++) { aL = mn(aL, xs(q)); bL = mx(bL, xs(q)); }
#pragma unroll
        for (int q = 29; q <= 36; q++) { aR = mn(aR, xs(q)); bR = mx(bR, xs(q)); }
        // odd ring of cell i (ll = 1,3,..,17) and even ring of cell i+1 (2,..,18)
        R ao = xs(1), bo = xs(1), ae = xs(2), be = xs(2);
#pragma unroll
        for (int q = 3; q <= 17; q += 2) { ao = mn(ao, xs(q)); bo = mx(bo, xs(q)); }
#pragma unroll
        for (int q = 4; q <= 18; q += 2) { ae = mn(ae, xs(q)); be = mx(be, xs(q)); }
        R qmin, qmax;
        if ((rxy_l < eps2) && (rxz_l < eps2)) { qmin = mn(a18, aL); qmax = mx(b18, bL); }
        else if ((x10 - x23) > R(0)) { qmax = b18; qmin = mn(ao, aL); }
        else { qmin = a18; qmax = mx(bo, bL); }
        R qml = mn(fabs(qmax - x9), fabs(qmin - x9));
        if ((rxy_r < eps2) && (rxz_r < eps2)) { qmin = mn(a18, aR); qmax = mx(b18, bR); }
        else if ((x32 - x9) > R(0)) { qmax = mx(be, bR); qmin = a18; }
        else { qmin = mn(ae, aR); qmax = b18; }
        R qmr = mn(fabs(qmax - x10), fabs(qmin - x10));
        R d1 = R(1) + rxy_l + rxz_l, d2 = R(1) + rxy_r + rxz_r;
        m.alp[0] = R(2) * mx(R(1), m.r[0]) / (fabs(dq) + eps) / d1 * qml;
        if (fabs(dq) <= eps) m.alp[1] = R(2) * mx(R(1), dqp / eps) * fabs(dqpp) / d2 * qmr;
        else m.alp[1] = R(2) * mx(R(1), R(1) / m.r[1]) * fabs(dqpp) / d2 * qmr;
        m.alp[0] = mx(R(1), mn(R(2), m.alp[0]));
        m.alp[1] = mx(R(1), mn(R(2), m.alp[1]));
    }
    xl = x9 + R(0.5) * limiter(nlim, m, 0) * dqm;
    xr = x10 - R(0.5) * limiter(nlim, m, 1) * dqp;
}

// positivity / fraction / turbulence clamps on the face states
template <class R> HD void muscl_clamps(int npv, R pref, const R *c9, const R *c10, R *xl, R *xr)
{
    xl[0] = mx(-pref + R(10), xl[0]);
    xr[0] = mx(-pref + R(10), xr[0]);
    if ((xl[5] < R(0)) || (xl[5] > R(1))) xl[5] = c9[5];
    if ((xr[5] < R(0)) || (xr[5] > R(1))) xr[5] = c10[5];
    if ((xl[6] < R(0)) || (xl[6] > R(1))) xl[6] = c9[6];
    if ((xr[6] < R(0)) || (xr[6] > R(1))) xr[6] = c10[6];
    if ((xl[5] + xl[6]) > R(1)) { xl[5] = c9[5]; xl[6] = c9[6]; }
    if ((xr[5] + xr[6]) > R(1)) { xr[5] = c10[5]; xr[6] = c10[6]; }
    if (npv > 7) {
        if ((xl[7] < R(0)) && (c9[7] > R(0))) xl[7] = c9[7];
        if ((xr[7] < R(0)) && (c10[7] > R(0))) xr[7] = c10[7];
        if ((xl[8] < R(0)) && (c9[8] > R(0))) xl[8] = c9[8];
        if ((xr[8] < R(0)) && (c10[8] > R(0))) xr[8] = c10[8];
    }
}

// ------------------------------------------------------------ inviscid flux
// pl/pr primitive face states (pl[0] gauge p), rl/rr densities, hl/hr static
// enthalpies, c2m the mixture c^2 at the Roe-averaged state.
template <class R> struct FaceState { R p[NPV]; R rho, h, c2; };

template <class R> HD void roe_flux(const Par &P, int npv, const R *n3, const FaceState<R> &L,
                                    const FaceState<R> &Rt, R c2avg, R *fx)
{
    R dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == R(0)) { for (int k = 0; k < NPV; k++) if (k < npv) fx[k] = 0; return; }
    R nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    const R *pl = L.p, *pr = Rt.p;
    R uull = nx * pl[1] + ny * pl[2] + nz * pl[3];
    R uurr = nx * pr[1] + ny * pr[2] + nz * pr[3];
    R sl = sqrt(L.rho), sr = sqrt(Rt.rho);
    R ravd = R(1) / (sl + sr);
    R ra[NPV];
#pragma unroll
    for (int k = 1; k < NPV; k++) if (k != 4 && k < npv) ra[k] = (sl * pl[k] + sr * pr[k]) * ravd;
    R hl = L.h + R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    R hr = Rt.h + R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    R raht = (sl * hl + sr * hr) * ravd;
    R u2 = ra[1] * ra[1] + ra[2] * ra[2] + ra[3] * ra[3];
    R rr = sqrt(L.rho * Rt.rho);
    R uuu = nx * ra[1] + ny * ra[2] + nz * ra[3];
    R sp = sndp2(P, c2avg, u2);
    R q = sp / c2avg, omq = R(1) - q;
    R uup = R(0.5) * (R(1) + q) * uuu;
    R ddd = R(0.5) * sqrt(omq * omq * (uuu * uuu) + R(4) * sp);
    R cs = R(0.5) * (fabs(uup + ddd) + fabs(uup - ddd));
    R ms = R(0.5) * (fabs(uup + ddd) - fabs(uup - ddd)) / ddd;
    R dpp = pr[0] - pl[0], dun = uurr - uull;
    R du = ms * dun + (cs - q * fabs(uuu) - R(0.5) * omq * uuu * ms) * dpp / rr / sp;
    R dp = ms * dpp + (cs - fabs(uuu) + R(0.5) * omq * uuu * ms) * rr * dun;
    R au = fabs(uuu), dur = du * rr;
    R ml = L.rho * uull, mr = Rt.rho * uurr;
    R pla = pl[0] + R(P.pref), pra = pr[0] + R(P.pref);
    fx[0] = R(0.5) * (ml + mr - (au * (Rt.rho - L.rho) + dur)) * dl;
    fx[1] = R(0.5) * (ml * pl[1] + nx * pla + mr * pr[1] + nx * pra - (au * (Rt.rho * pr[1] - L.rho * pl[1]) + dur * ra[1] + dp * nx)) * dl;
    fx[2] = R(0.5) * (ml * pl[2] + ny * pla + mr * pr[2] + ny * pra - (au * (Rt.rho * pr[2] - L.rho * pl[2]) + dur * ra[2] + dp * ny)) * dl;
    fx[3] = R(0.5) * (ml * pl[3] + nz * pla + mr * pr[3] + nz * pra - (au * (Rt.rho * pr[3] - L.rho * pl[3]) + dur * ra[3] + dp * nz)) * dl;
    fx[4] = R(0.5) * (ml * hl + mr * hr - (au * (Rt.rho * hr - pr[0] - L.rho * hl + pl[0]) + dur * raht + dp * uuu)) * dl;
#pragma unroll
    for (int k = 5; k < NPV; k++)
        if (k < npv) fx[k] = R(0.5) * (ml * pl[k] + mr * pr[k] - (au * (Rt.rho * pr[k] - L.rho * pl[k]) + dur * ra[k])) * dl;
}

template <class R> HD void ausm_flux(const Par &P, int npv, const R *n3, const FaceState<R> &L,
                                     const FaceState<R> &Rt, R c2avg, R *fx)
{
    const R beta = R(0.125), kp = R(0.25), ku = R(1);
    R dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == R(0)) { for (int k = 0; k < NPV; k++) if (k < npv) fx[k] = 0; return; }
    R nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    const R *pl = L.p, *pr = Rt.p;
    R uull = nx * pl[1] + ny * pl[2] + nz * pl[3];
    R uurr = nx * pr[1] + ny * pr[2] + nz * pr[3];
    R sl = sqrt(L.rho), sr = sqrt(Rt.rho);
    R ravd = R(1) / (sl + sr);
    R ua = (sl * pl[1] + sr * pr[1]) * ravd, va = (sl * pl[2] + sr * pr[2]) * ravd, wa = (sl * pl[3] + sr * pr[3]) * ravd;
    R rr = sqrt(L.rho * Rt.rho);
    R amid = sqrt(c2avg);
    R zmr = uurr / amid, zml = uull / amid;
    R un = nx * ua + ny * va + nz * wa;
    R am2 = un * un;
    R am2r = sndp2(P, c2avg, am2) / c2avg;
    R fmid = sqrt(am2r) * (R(2) - sqrt(am2r));
    R alpha = R(0.1875) * (R(-4) + R(5) * (fmid * fmid));
    alpha = mx(mn(alpha, R(0.1875)), R(-0.75));
    R zmmr, pmr, zmpl, ppl;
    if (fabs(zmr) >= R(1)) {
        zmmr = R(0.5) * (zmr - fabs(zmr));
        pmr = R(0.5) * (R(1) - zmr / fabs(zmr));
    } else {
        R t1 = (zmr - R(1)) * (zmr - R(1)), t2 = (zmr * zmr - R(1)) * (zmr * zmr - R(1));
        zmmr = -(R(0.25) * t1) - beta * t2;
        pmr = R(0.25) * t1 * (R(2) + zmr) - alpha * zmr * t2;
    }
    if (fabs(zml) >= R(1)) {
        zmpl = R(0.5) * (zml + fabs(zml));
        ppl = R(0.5) * (R(1) + zml / fabs(zml));
    } else {
        R t1 = (zml + R(1)) * (zml + R(1)), t2 = (zml * zml - R(1)) * (zml * zml - R(1));
        zmpl = R(0.25) * t1 + beta * t2;
        ppl = R(0.25) * t1 * (R(2) - zml) + alpha * zml * t2;
    }
    R zmid = zmpl + zmmr - kp * mx(R(1) - am2 / c2avg, R(0)) * (pr[0] - pl[0]) / rr / c2avg / fmid;
    R pmid = ppl * (pl[0] + R(P.pref)) + pmr * (pr[0] + R(P.pref)) - R(2) * ku * ppl * pmr * rr * fmid * amid * (uurr - uull);
    R ml = (zmid > R(0)) ? zmid * amid * L.rho : R(0);
    R mr = (zmid > R(0)) ? R(0) : zmid * amid * Rt.rho;
    R hl = L.h + R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    R hr = Rt.h + R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    fx[0] = (ml + mr) * dl;
    fx[1] = (ml * pl[1] + mr * pr[1] + pmid * nx) * dl;
    fx[2] = (ml * pl[2] + mr * pr[2] + pmid * ny) * dl;
    fx[3] = (ml * pl[3] + mr * pr[3] + pmid * nz) * dl;
    fx[4] = (ml * hl + mr * hr) * dl;
#pragma unroll
    for (int k = 5; k < NPV; k++) if (k < npv) fx[k] = (ml * pl[k] + mr * pr[k]) * dl;
}

// Shock-sensor pressures sd[18] (gauge, the face's 18-point stencil), lifted
// by p_ref and a tenth of the smaller rho*c^2; the smallest of the ratios the
// Roe-M and AUSMPW+ fluxes compare (src/flux_module.f90:375-386, 499-508)
template <class R> HD void press_ratios(const Par &P, const FaceState<R> &L, const FaceState<R> &Rt, const R *sd,
                                        R &m910, R &mall)
{
    R lift = R(0.1) * mn(L.rho * L.c2, Rt.rho * Rt.c2);
    R s[18];
#pragma unroll
    for (int k = 0; k < 18; k++) s[k] = sd[k] + R(P.pref) + lift;
    // 1-based pairs of the reference: (9,10) (11,9) (9,7) (9,3) (9,15) (10,12) (10,8) (4,10) (10,16)
    const int a[9] = {8, 10, 8, 8, 8, 9, 9, 3, 9}, b[9] = {9, 8, 6, 2, 14, 11, 7, 9, 15};
    m910 = mn(s[8] / s[9], s[9] / s[8]);
    mall = m910;
#pragma unroll
    for (int q = 1; q < 9; q++) mall = mn(mall, mn(s[a[q]] / s[b[q]], s[b[q]] / s[a[q]]));
}

// Roe-M (nscheme 2), src/flux_module.f90:278-418
template <class R> HD void roem_flux(const Par &P, int npv, const R *n3, const FaceState<R> &L, const FaceState<R> &Rt,
                                     R c2avg, const R *sd, R *fx)
{
    R dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == R(0)) { for (int k = 0; k < NPV; k++) if (k < npv) fx[k] = 0; return; }
    R nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    const R *pl = L.p, *pr = Rt.p;
    R uull = nx * pl[1] + ny * pl[2] + nz * pl[3];
    R uurr = nx * pr[1] + ny * pr[2] + nz * pr[3];
    R sl = sqrt(L.rho), sr = sqrt(Rt.rho);
    R ravd = R(1) / (sl + sr);
    R ra[NPV];
#pragma unroll
    for (int k = 1; k < NPV; k++) if (k != 4 && k < npv) ra[k] = (sl * pl[k] + sr * pr[k]) * ravd;
    R hl = L.h + R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    R hr = Rt.h + R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    R raht = (sl * hl + sr * hr) * ravd;
    R u2 = ra[1] * ra[1] + ra[2] * ra[2] + ra[3] * ra[3];
    R rr = sqrt(L.rho * Rt.rho);
    R uuu = nx * ra[1] + ny * ra[2] + nz * ra[3];
    R am2du = sndp2_du(P, c2avg, u2) / c2avg;
    R am2 = sndp2(P, c2avg, u2) / c2avg;
    R fmid = sqrt(am2) * (R(2) - sqrt(am2));
    R fmdu = sqrt(am2du) * (R(2) - sqrt(am2du));
    R amid = sqrt(c2avg);
    R b1 = mx(mx(R(0), uuu + amid), uurr + amid);
    R b2 = mn(mn(R(0), uuu - amid), uull - amid);
    R b1du = mx(mx(R(0), uuu + fmdu * amid), uurr + fmdu * amid);
    R b2du = mn(mn(R(0), uuu - fmdu * amid), uull - fmdu * amid);
    R sgn = signbit(uuu) ? R(-1) : R(1);
    R mt = sgn * mn(fabs(uuu) / amid, R(1));
    R mtdu = sgn * mn(fabs(uuu) / (fmdu * amid), R(1));
    R m910, mall;
    press_ratios<R>(P, L, Rt, sd, m910, mall);
    R ff = R(1) - mall, gg = R(1) - m910;
    ff = uuu != R(0) ? pow(fabs(uuu) / amid, ff) : R(1);
    gg = uuu != R(0) ? pow(fabs(uuu) / amid, gg) : R(1);
    R add = R(1) + fabs(mt), aaa = R(1) + fabs(mtdu);
    R dq[NPV], fl[NPV], fr[NPV], bdq[NPV];
    dq[0] = Rt.rho - L.rho;
    fl[0] = L.rho * uull;
    fr[0] = Rt.rho * uurr;
#pragma unroll
    for (int k = 1; k < NPV; k++) if (k < npv) {
        R ql = k == 4 ? hl : pl[k], qr = k == 4 ? hr : pr[k];
        dq[k] = Rt.rho * qr - L.rho * ql;
        fl[k] = L.rho * uull * ql;
        fr[k] = Rt.rho * uurr * qr;
    }
    fl[1] = fl[1] + nx * (pl[0] + R(P.pref)); fr[1] = fr[1] + nx * (pr[0] + R(P.pref));
    fl[2] = fl[2] + ny * (pl[0] + R(P.pref)); fr[2] = fr[2] + ny * (pr[0] + R(P.pref));
    fl[3] = fl[3] + nz * (pl[0] + R(P.pref)); fr[3] = fr[3] + nz * (pr[0] + R(P.pref));
    bdq[0] = dq[0] - ff * (pr[0] - pl[0]) / c2avg / fmid;
#pragma unroll
    for (int k = 1; k < NPV; k++) if (k < npv)
        bdq[k] = k == 4 ? bdq[0] * raht + rr * (hr - hl) : bdq[0] * ra[k] + rr * (pr[k] - pl[k]);
#pragma unroll
    for (int k = 0; k < NPV; k++) if (k < npv)
        fx[k] = (b1 * fl[k] - b2 * fr[k] + b1 * b2 * (dq[k] - gg * bdq[k] / add)) / (b1 - b2) * dl;
    R du = gg * b1du * b2du * rr * (uurr - uull);
    fx[1] = fx[1] + du * nx / aaa / (b1du - b2du) * dl;
    fx[2] = fx[2] + du * ny / aaa / (b1du - b2du) * dl;
    fx[3] = fx[3] + du * nz / aaa / (b1du - b2du) * dl;
}

// AUSMPW+ (nscheme 3), src/flux_module.f90:420-546
template <class R> HD void ausmpwp_flux(const Par &P, int npv, const R *n3, const FaceState<R> &L,
                                        const FaceState<R> &Rt, R c2avg, const R *sd, R *fx)
{
    const R beta = R(0.125), ku = R(1);
    R dl = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    if (dl == R(0)) { for (int k = 0; k < NPV; k++) if (k < npv) fx[k] = 0; return; }
    R nx = n3[0] / dl, ny = n3[1] / dl, nz = n3[2] / dl;
    const R *pl = L.p, *pr = Rt.p;
    R uull = nx * pl[1] + ny * pl[2] + nz * pl[3];
    R uurr = nx * pr[1] + ny * pr[2] + nz * pr[3];
    R sl = sqrt(L.rho), sr = sqrt(Rt.rho);
    R ravd = R(1) / (sl + sr);
    R ua = (sl * pl[1] + sr * pr[1]) * ravd, va = (sl * pl[2] + sr * pr[2]) * ravd, wa = (sl * pl[3] + sr * pr[3]) * ravd;
    R rr = sqrt(L.rho * Rt.rho);
    R amid = sqrt(c2avg);
    R zmr = uurr / amid, zml = uull / amid;
    R am2 = ua * ua + va * va + wa * wa;
    R am2r1 = sndp2_du(P, c2avg, am2) / c2avg;
    R am2r = sndp2(P, c2avg, am2) / c2avg;
    R fmid = sqrt(am2r) * (R(2) - sqrt(am2r));
    R fmid1 = sqrt(am2r1) * (R(2) - sqrt(am2r1));
    R alpha = R(0.1875) * (R(-4) + R(5) * (fmid1 * fmid1));
    alpha = mx(mn(alpha, R(0.1875)), R(-0.75));
    R zmmr, pmr, zmpl, ppl;
    if (fabs(zmr) >= R(1)) {
        zmmr = R(0.5) * (zmr - fabs(zmr));
        pmr = R(0.5) * (R(1) - zmr / fabs(zmr));
    } else {
        R t1 = (zmr - R(1)) * (zmr - R(1)), t2 = (zmr * zmr - R(1)) * (zmr * zmr - R(1));
        zmmr = -(R(0.25) * t1) - beta * t2;
        pmr = R(0.25) * t1 * (R(2) + zmr) - alpha * zmr * t2;
    }
    if (fabs(zml) >= R(1)) {
        zmpl = R(0.5) * (zml + fabs(zml));
        ppl = R(0.5) * (R(1) + zml / fabs(zml));
    } else {
        R t1 = (zml + R(1)) * (zml + R(1)), t2 = (zml * zml - R(1)) * (zml * zml - R(1));
        zmpl = R(0.25) * t1 + beta * t2;
        ppl = R(0.25) * t1 * (R(2) - zml) + alpha * zml * t2;
    }
    R zmid = zmpl + zmmr;
    R pmid = ppl * (pl[0] + R(P.pref)) + pmr * (pr[0] + R(P.pref)) - R(2) * ku * ppl * pmr * rr * fmid1 * amid * (uurr - uull);
    R m910, mall;
    press_ratios<R>(P, L, Rt, sd, m910, mall);
    R ww1 = R(1) - m910 * m910 * m910;
    R ww2 = R(1) - mall * mall;
    R ww = mx(ww1, ww2);
    R pmt = rr * c2avg;
    R pwl = ((pl[0] + R(P.pref) + rr * c2avg) / pmt - R(1)) * (R(1) - ww) / fmid;
    R pwr = ((pr[0] + R(P.pref) + rr * c2avg) / pmt - R(1)) * (R(1) - ww) / fmid;
    R zmpl1, zmmr1;
    if (zmid >= R(0)) {
        pwl = pwl * rr / L.rho;
        pwr = pwr * rr / L.rho;
        zmpl1 = zmpl + zmmr * ((R(1) - ww) * (R(1) + pwr) - pwl);
        zmmr1 = zmmr * ww * (R(1) + pwr);
    } else {
        pwl = pwl * rr / Rt.rho;
        pwr = pwr * rr / Rt.rho;
        zmpl1 = zmpl * ww * (R(1) + pwl);
        zmmr1 = zmmr + zmpl * ((R(1) - ww) * (R(1) + pwl) - pwr);
    }
    R hl = L.h + R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    R hr = Rt.h + R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    fx[0] = (zmpl1 * amid * L.rho + zmmr1 * amid * Rt.rho) * dl;
    fx[1] = (zmpl1 * amid * L.rho * pl[1] + zmmr1 * amid * Rt.rho * pr[1] + pmid * nx) * dl;
    fx[2] = (zmpl1 * amid * L.rho * pl[2] + zmmr1 * amid * Rt.rho * pr[2] + pmid * ny) * dl;
    fx[3] = (zmpl1 * amid * L.rho * pl[3] + zmmr1 * amid * Rt.rho * pr[3] + pmid * nz) * dl;
    fx[4] = (zmpl1 * amid * L.rho * hl + zmmr1 * amid * Rt.rho * hr) * dl;
#pragma unroll
    for (int k = 5; k < NPV; k++) if (k < npv) fx[k] = (zmpl1 * amid * L.rho * pl[k] + zmmr1 * amid * Rt.rho * pr[k]) * dl;
}

// ------------------------------------------------------------ one face
// Reconstruction of every variable at a face from its 38-point stencil,
// xs(ll, n) = stencil value at Fortran index ll of variable n
// (src/rhs_module.f90:288-312), plus the normal / transverse differences of
// u, v, w, T, k, omega the viscous flux uses (src/vsflux_module.f90:156-180).
template <class R, class X>
HD void face_recon(const Par &P, int npv, X xs, R *c9, R *c10, R *xl, R *xr, R g3[6][3])
{
    const bool visc = (P.iturb >= -2);
    int vq = 0;
    for (int n = 0; n < NPV; n++) if (n < npv) {
        auto xn = [&](int ll) -> R { return xs(ll, n); };
        c9[n] = xn(9);
        c10[n] = xn(10);
        if (P.nmuscl > 0) muscl_var<R>(P.nlim, P.nmuscl == 2, xn, xl[n], xr[n]);
        else { xl[n] = c9[n]; xr[n] = c10[n]; }
        if (visc && ((n >= 1 && n <= 4) || n >= 7)) {
            g3[vq][0] = xn(10) - xn(9);
            g3[vq][1] = R(0.0625) * (-xn(1) - xn(2) + xn(5) + xn(6) + R(2) * (-xn(7) - xn(8) + xn(11) + xn(12))
                                     - xn(13) - xn(14) + xn(17) + xn(18));
            g3[vq][2] = R(0.0625) * (-xn(1) - xn(2) + xn(13) + xn(14) + R(2) * (-xn(3) - xn(4) + xn(15) + xn(16))
                                     - xn(5) - xn(6) + xn(17) + xn(18));
            vq++;
        }
    }
}

// Inviscid flux of a face from the reconstructed states: clamps, face EOS,
// Roe-averaged c^2 and the Roe / AUSM+-up flux (src/rhs_module.f90:314-337).
// rh9/rh10 = {rho, h} of the two cells (first-order faces read the cells').
// SENSOR: Roe-M / AUSMPW+ (they read the 18 stencil pressures sd and the
// face sound speeds); otherwise Roe / AUSM+-up.
template <class R, bool SENSOR = false>
HD void inviscid_face(const Par &P, int npv, const R *n3, const R *c9, const R *c10, R *xl, R *xr,
                      const R *rh9, const R *rh10, R *fx, const R *sd = nullptr)
{
    FaceState<R> L, Rt;
    constexpr bool sensor = SENSOR;
    if (P.nmuscl > 0) {
        muscl_clamps<R>(npv, R(P.pref), c9, c10, xl, xr);
        face_rho_h<R>(xl[0] + R(P.pref), xl[4], xl[5], xl[6], L.rho, L.h);
        face_rho_h<R>(xr[0] + R(P.pref), xr[4], xr[5], xr[6], Rt.rho, Rt.h);
        if (sensor) {
            L.c2 = face_c2<R>(xl[0] + R(P.pref), xl[4], xl[5], xl[6]);
            Rt.c2 = face_c2<R>(xr[0] + R(P.pref), xr[4], xr[5], xr[6]);
        }
    } else {
        L.rho = rh9[0]; L.h = rh9[1];
        Rt.rho = rh10[0]; Rt.h = rh10[1];
        if (sensor) { L.c2 = rh9[2]; Rt.c2 = rh10[2]; }
    }
    for (int n = 0; n < NPV; n++) if (n < npv) { L.p[n] = xl[n]; Rt.p[n] = xr[n]; }
    R sl = sqrt(L.rho), sr = sqrt(Rt.rho), rd = R(1) / (sl + sr);
    R ya1 = (sl * xl[5] + sr * xr[5]) * rd, ya2 = (sl * xl[6] + sr * xr[6]) * rd;
    R c2a = face_c2<R>(R(0.5) * (xl[0] + xr[0]) + R(P.pref), R(0.5) * (xl[4] + xr[4]), ya1, ya2);
    if constexpr (SENSOR) {
        if (P.nscheme == 2) roem_flux<R>(P, npv, n3, L, Rt, c2a, sd, fx);
        else ausmpwp_flux<R>(P, npv, n3, L, Rt, c2a, sd, fx);
    } else {
        if (P.nscheme == 1) roe_flux<R>(P, npv, n3, L, Rt, c2a, fx);
        else ausm_flux<R>(P, npv, n3, L, Rt, c2a, fx);
    }
}

// ------------------------------------------------------------ Jacobian / getx
// Cell data the split Jacobian and the diagonal inverse read.
struct CellJ {
    double pv[NPV];
    double d[ND];
    double mut, vol;
};

HD double eigenvis(const Par &P, const double *d, double mut)
{
    if (P.iturb == -3) return 0.0;
    double den = d[D_RT] + d[D_HT] * d[D_RP] * d[D_RHO] - d[D_HP] * d[D_RT] * d[D_RHO];
    if (P.iturb == -2) return mx(d[D_RP] * d[D_KAP] / den, 4.0 / 3.0 * d[D_MU] / d[D_RHO]);
    return mx(d[D_RP] * (d[D_KAP] + d[D_HT] * mut / 0.9) / den, 4.0 / 3.0 * (d[D_MU] + mut) / d[D_RHO]);
}

// The split flux Jacobian a(n,l) of a cell through one of its faces
// (src/jacobian_module.f90:312-402), applied as y = y + sgn * A x row by row
// over l in the reference's order (src/update_module.f90:507-544).  The
// cell-only terms (JacCell) and the face terms (JacFace, for sign +1; the
// sign flips lam and vis exactly) are computed once per iteration; entries
// that are identically zero are skipped (adding a zero product leaves y
// unchanged).
struct JacCell { double beta, h; };
struct JacFace { double u, lam, vis; };

HD JacCell jac_cell(const Par &P, const CellJ &c)
{
    const double *pv = c.pv, *d = c.d;
    double sp = sndp2(P, d[D_C2], pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
    JacCell j;
    j.beta = 1.0 / sp - 1.0 / d[D_C2] + d[D_RP];
    j.h = d[D_H] + 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
    return j;
}

HD JacFace jac_face(const Par &P, const CellJ &c, const double *n3)
{
    const double *pv = c.pv, *d = c.d;
    const double cappa = 1.05;
    double ds = n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2];
    double u = n3[0] * pv[1] + n3[1] * pv[2] + n3[2] * pv[3];
    double sp = sndp2(P, d[D_C2], pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
    double up = 0.5 * (1.0 + sp / d[D_C2]) * u;
    double omq = 1.0 - sp / d[D_C2];
    double dd = 0.5 * sqrt((u * u) * (omq * omq) + 4.0 * sp * ds);
    JacFace f;
    f.u = u;
    f.lam = cappa * (fabs(up) + dd);
    f.vis = 2.0 * eigenvis(P, d, c.mut) * ds / c.vol;
    return f;
}

// What row n of a Jacobian reads: the cell's pv(n), velocity, density
// derivatives and enthalpy derivatives, the face vector and the
// precomputed cell/face terms.
struct JacNb {
    double pvn, pv1, pv2, pv3;
    double rho, rp, rt, ry1, ry2, hp, ht, hy1, hy2;
    double n3[3];
    double n3n;       // n3[n-1] for rows 1..3 (rows are lanes: no run-time register indexing)
    JacCell jc;
    JacFace jf;
};

HD JacNb jac_nb(const double *pv, const double *d, const double *n3, const JacCell &jc, const JacFace &jf, int n, int npv)
{
    JacNb q;
    q.pvn = (n < npv) ? pv[n] : 0.0;
    q.pv1 = pv[1]; q.pv2 = pv[2]; q.pv3 = pv[3];
    q.rho = d[D_RHO]; q.rp = d[D_RP]; q.rt = d[D_RT]; q.ry1 = d[D_RY1]; q.ry2 = d[D_RY2];
    q.hp = d[D_HP]; q.ht = d[D_HT]; q.hy1 = d[D_HY1]; q.hy2 = d[D_HY2];
    q.n3[0] = n3[0]; q.n3[1] = n3[1]; q.n3[2] = n3[2];
    q.n3n = (n >= 1 && n <= 3) ? n3[n - 1] : 0.0;
    q.jc = jc;
    q.jf = jf;
    return q;
}

// row n of y += sgn * A x; xn = x[n]
HD void jac_row(int npv, const JacNb &q, int sign, int n, const double *x, double xn, double &y, double sgn)
{
    const double lam = sign > 0 ? q.jf.lam : -q.jf.lam, vis = sign > 0 ? q.jf.vis : -q.jf.vis;
    const double u = q.jf.u;
    double r1[7];
    r1[0] = u * q.rp + lam * q.jc.beta + vis * q.rp;
    r1[1] = q.n3[0] * q.rho;
    r1[2] = q.n3[1] * q.rho;
    r1[3] = q.n3[2] * q.rho;
    r1[4] = u * q.rt + lam * q.rt + vis * q.rt;
    r1[5] = u * q.ry1 + lam * q.ry1 + vis * q.ry1;
    r1[6] = u * q.ry2 + lam * q.ry2 + vis * q.ry2;
    const double uuu = u + lam + vis;
    const double ru = q.rho * uuu;
    auto acc = [&](double a, int l) {
        double t = (0.5 * a) * x[l];
        y = (sgn > 0.0) ? y + t : y - t;
    };
    if (n == 0) {
#pragma unroll
        for (int l = 0; l < 7; l++) acc(r1[l], l);
    } else if (n <= 3) {
#pragma unroll
        for (int l = 0; l < 7; l++) {
            double a = q.pvn * r1[l];
            if (l == 0) a = q.pvn * r1[0] + q.n3n;
            if (l == n) a = q.pvn * r1[n] + ru;
            acc(a, l);
        }
    } else if (n == 4) {
        const double hy[7] = {q.hp, q.pv1, q.pv2, q.pv3, q.ht, q.hy1, q.hy2};
#pragma unroll
        for (int l = 0; l < 7; l++) {
            double a = ru * hy[l] + q.jc.h * r1[l];
            if (l == 0) a = a - lam - vis;
            acc(a, l);
        }
    } else if (n < npv) {
#pragma unroll
        for (int l = 0; l < 7; l++) {
            double a = q.pvn * r1[l];
            if (l == n) a = q.pvn * r1[n] + ru;
            acc(a, l);
        }
        if (n >= 7) {
            double t = (0.5 * (uuu * q.rho)) * xn;
            y = (sgn > 0.0) ? y + t : y - t;
        }
    }
}

// Row n of the same product as coefficients, formed without branching on n
// (the rows of a cell sit on neighbouring lanes): y = y + c[l] x[l] over
// l = 0..6 in order, then y = y + c[7] x[n], repeats jac_row's roundings.
// Every a(n,l) is built from jac_row's products in jac_row's order; the
// terms a row does not have enter as exact zeros (P r1 + 0 hy, a + 0), the
// factor 0.5 and the sweep's sign are folded in exactly (c = +-0.5 a), and
// rows n >= npv get zero coefficients.  Only the sign of a zero sum can
// differ from jac_row.
HD void jac_row_coef(int npv, const JacNb &q, int sign, int n, double sgn, double *c)
{
    const double lam = sign > 0 ? q.jf.lam : -q.jf.lam, vis = sign > 0 ? q.jf.vis : -q.jf.vis;
    const double u = q.jf.u;
    double r1[7];
    r1[0] = u * q.rp + lam * q.jc.beta + vis * q.rp;
    r1[1] = q.n3[0] * q.rho;
    r1[2] = q.n3[1] * q.rho;
    r1[3] = q.n3[2] * q.rho;
    r1[4] = u * q.rt + lam * q.rt + vis * q.rt;
    r1[5] = u * q.ry1 + lam * q.ry1 + vis * q.ry1;
    r1[6] = u * q.ry2 + lam * q.ry2 + vis * q.ry2;
    const double uuu = u + lam + vis;
    const double ru = q.rho * uuu;
    const double hy[7] = {q.hp, q.pv1, q.pv2, q.pv3, q.ht, q.hy1, q.hy2};
    const bool mom = n >= 1 && n <= 3, en = n == 4, row = n < npv;
    const double P = n == 0 ? 1.0 : (en ? q.jc.h : q.pvn);
    const double Q = en ? ru : 0.0;
    const double hs = sgn > 0.0 ? 0.5 : -0.5;
#pragma unroll
    for (int l = 0; l < 7; l++) {
        double a = P * r1[l] + Q * hy[l];
        const double e1 = l == 0 ? (mom ? q.n3n : (en ? -lam : 0.0)) : ((l == n && !en) ? ru : 0.0);
        a = a + e1;
        a = a + ((en && l == 0) ? -vis : 0.0);
        c[l] = row ? hs * a : 0.0;
    }
    c[7] = (row && n >= 7) ? hs * (uuu * q.rho) : 0.0;
}

HD void jac_row_apply(const double *c, const double *x, double xn, double &y)
{
#pragma unroll
    for (int l = 0; l < 7; l++) y = y + c[l] * x[l];
    y = y + c[7] * xn;
}

HD void jac_apply(const Par &P, int npv, const CellJ &c, const double *n3, int sign, const double *x, double *y, double sgn)
{
    JacCell jc = jac_cell(P, c);
    JacFace jf = jac_face(P, c, n3);
    for (int n = 0; n < NPV; n++) if (n < npv) jac_row(npv, jac_nb(c.pv, c.d, n3, jc, jf, n, npv), sign, n, x, x[n], y[n], sgn);
}

// Closed-form inverse of the preconditioned diagonal block with the Merkle
// (icav) and SST (itt) couplings (src/lhs_module.f90:346-457), split into the
// part that depends only on the cell (GxPre, computed once per iteration in a
// parallel pass) and the part linear in the right-hand side (getx_apply, run
// inside the sweeps).  Products are grouped exactly as the reference groups
// them, so the split changes no rounding.
enum {
    GX_A = 0, GX_IB1, GX_MM, GX_SS, GX_AAA1, GX_AAA2, GX_CROSS, GX_CROSS2, GX_CROSS1, GX_SSHM, GX_RRR,
    GX_SSY1, GX_SSY2, GX_C1, GX_C2, GX_C3, GX_C4C1, GX_RB1, GX_CP5, GX_C4IB1S1, GX_C3IB1, GX_S1P5,
    GX_C4SSIB1, GX_C2IB1, GX_T2, GX_T3, GX_RBT4, GX_RBT1, GX_IT, NGX
};

HD void getx_prep(const Par &P, int npv, const CellJ &c, const double f6[6][3], double dt,
                  const double *cc, const double *tt4, double *g)
{
    const double *pv = c.pv, *d = c.d;
    double v2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
    double sp = sndp2(P, d[D_C2], v2);
    double q = sp / d[D_C2], omq = 1.0 - q;
    double eig = 0.0, msum = 0.0;
    for (int dd = 0; dd < 3; dd++) {
        double o0 = 0.5 * (f6[2 * dd][0] + f6[2 * dd + 1][0]);
        double o1 = 0.5 * (f6[2 * dd][1] + f6[2 * dd + 1][1]);
        double o2 = 0.5 * (f6[2 * dd][2] + f6[2 * dd + 1][2]);
        double ds = o0 * o0 + o1 * o1 + o2 * o2;
        double u = o0 * pv[1] + o1 * pv[2] + o2 * pv[3];
        double up = 0.5 * (1.0 + q) * u;
        double e = 0.5 * sqrt((u * u) * (omq * omq) + 4.0 * sp * ds);
        eig += 1.05 * (fabs(up) + e);
        for (int m = 0; m < 3; m++) {   // src/lhs_module.f90:525-527: nine terms in sequence
            double qq = f6[2 * dd][m] + f6[2 * dd + 1][m];
            msum = (dd == 0 && m == 0) ? 0.25 * (qq * qq) : msum + 0.25 * (qq * qq);
        }
    }
    double ev = (P.iturb == -3) ? 0.0 : eigenvis(P, d, c.mut) * msum;
    double a = 1.0 / (c.vol / dt + eig);
    double b = ((double)P.nsteady * P.bdf_b * c.vol / P.dt_phy + 2.0 * ev / c.vol) * a;
    double c1 = cc[0] * c.vol * a, c2 = cc[1] * c.vol * a, c3 = cc[2] * c.vol * a, c4 = cc[3] * c.vol * a;
    double b1 = 1.0 + b;
    double rb = b1 * d[D_RHO];
    double i_b1 = 1.0 / rb;
    double ss = d[D_RHO] * d[D_HT] * (1.0 / sp + b / d[D_C2]);
    double rr = 1.0 / sp - 1.0 / d[D_C2] + b1 * d[D_RP];
    double hst = d[D_H] + 0.5 * v2 - v2;
    double cross = d[D_HY2] * d[D_RY1] - d[D_HY1] * d[D_RY2];
    double cross1 = d[D_HY1] * d[D_RT] - d[D_HT] * d[D_RY1];
    double cross2 = d[D_HY2] * d[D_RT] - d[D_HT] * d[D_RY2];
    double dhdp1 = 1.0 - d[D_HP] * d[D_RHO];
    double ssy1 = d[D_HY1] * d[D_RHO] * rr + b1 * d[D_RY1] * dhdp1;
    double ssy2 = d[D_HY2] * d[D_RHO] * rr + b1 * d[D_RY2] * dhdp1;
    double ssh = d[D_RHO] * rr * hst - b1 * dhdp1 * d[D_RHO];
    g[GX_A] = a;
    g[GX_IB1] = i_b1;
    g[GX_MM] = 1.0 / (rb * (ss + c1 * cross1) - c2 * ssy1 + c3 * ss);
    g[GX_SS] = ss;
    g[GX_AAA1] = d[D_RY1] * hst + d[D_HY1] * d[D_RHO] - cross * pv[6];
    g[GX_AAA2] = d[D_RT] * hst + d[D_HT] * d[D_RHO] - cross2 * pv[6];
    g[GX_CROSS] = cross;
    g[GX_CROSS2] = cross2;
    g[GX_CROSS1] = cross1;
    g[GX_SSHM] = ssh - ssy2 * pv[6];
    g[GX_RRR] = d[D_RHO] * rr;
    g[GX_SSY1] = ssy1;
    g[GX_SSY2] = ssy2;
    g[GX_C1] = c1;
    g[GX_C2] = c2;
    g[GX_C3] = c3;
    g[GX_C4C1] = c4 * cross1;
    g[GX_RB1] = d[D_RHO] * b1;
    g[GX_CP5] = cross1 * pv[5];
    g[GX_C4IB1S1] = c4 * i_b1 * ssy1;
    g[GX_C3IB1] = c3 * i_b1;
    g[GX_S1P5] = ssy1 * pv[5];
    g[GX_C4SSIB1] = c4 * ss * i_b1;
    g[GX_C2IB1] = c2 * i_b1;
    if (npv > 7) {
        double t1 = tt4[0] * c.vol * a, t2 = tt4[1] * c.vol * a, t3 = tt4[2] * c.vol * a, t4 = tt4[3] * c.vol * a;
        g[GX_T2] = t2;
        g[GX_T3] = t3;
        g[GX_RBT4] = rb + t4;
        g[GX_RBT1] = rb + t1;
        g[GX_IT] = 1.0 / (t2 * t3 - (t1 + rb) * (t4 + rb));
    } else {
        g[GX_T2] = g[GX_T3] = g[GX_RBT4] = g[GX_RBT1] = g[GX_IT] = 0.0;
    }
}

// Component n of x = D^-1 res; G(q) returns the GxPre entry q.
template <class F>
HD double getx_apply(F G, const double *pv, double d_ry1, double d_rt, const double *res, int n, double resn, double pvn)
{
    const double a = G(GX_A);
    if (n >= 1 && n <= 3) return (resn - res[0] * pvn) * G(GX_IB1) * a;
    double r6 = res[6] - res[0] * pv[6];
    if (n == 6) return r6 * G(GX_IB1) * a;
    if (n >= 7) {
        double rk = res[7] - res[0] * pv[7], ro = res[8] - res[0] * pv[8];
        if (n == 7) return (ro * G(GX_T2) - G(GX_RBT4) * rk) * G(GX_IT) * a;
        return (rk * G(GX_T3) - G(GX_RBT1) * ro) * G(GX_IT) * a;
    }
    double uuu = res[1] * pv[1] + res[2] * pv[2] + res[3] * pv[3] - res[4];
    double co1 = res[0] * G(GX_AAA1) + d_ry1 * uuu + G(GX_CROSS) * res[6];
    double co2 = res[0] * G(GX_AAA2) + d_rt * uuu + G(GX_CROSS2) * res[6];
    double co3 = res[0] * G(GX_SSHM) + G(GX_RRR) * uuu + res[6] * G(GX_SSY2);
    double mm = G(GX_MM);
    if (n == 0)
        return (G(GX_C4C1) * (-r6) - G(GX_C2) * co1 + G(GX_C3) * co2
                + G(GX_RB1) * (co2 + G(GX_CROSS1) * res[5] - G(GX_CP5) * res[0])) * mm * a;
    if (n == 4)
        return (G(GX_C4IB1S1) * r6 + G(GX_C1) * co1 - G(GX_C3IB1) * co3
                - (co3 - G(GX_S1P5) * res[0] + res[5] * G(GX_SSY1))) * mm * a;
    return (G(GX_SS) * (res[5] - res[0] * pv[5]) - G(GX_C4SSIB1) * r6 - G(GX_C1) * co2 + G(GX_C2IB1) * co3) * mm * a;
}

// a where c holds, else b, bit for bit and without a branch: the mask is
// opaque to the compiler on the device, so it cannot sink the work behind a
// or b into a branch on c (the rows of a cell sit on neighbouring lanes).
HD double bsel(bool c, double a, double b)
{
    unsigned long long m = c ? ~0ull : 0ull;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(m));
#endif
    return __builtin_bit_cast(double, (__builtin_bit_cast(unsigned long long, a) & m) |
                                      (__builtin_bit_cast(unsigned long long, b) & ~m));
}

// getx_apply for any row without branching on n: every row's value is
// formed with getx_apply's operations in getx_apply's order, then row n's is
// selected (the rows of a cell sit on neighbouring lanes).
template <class F>
HD double getx_apply_sel(F G, const double *pv, double d_ry1, double d_rt, const double *res, int n, double resn, double pvn)
{
    const double a = G(GX_A);
    const double v1 = (resn - res[0] * pvn) * G(GX_IB1) * a;          // rows 1..3, and row 6 (resn = res[6], pvn = pv[6])
    const double r6 = res[6] - res[0] * pv[6];
    const double rk = res[7] - res[0] * pv[7], ro = res[8] - res[0] * pv[8];
    const double v7 = (ro * G(GX_T2) - G(GX_RBT4) * rk) * G(GX_IT) * a;
    const double v8 = (rk * G(GX_T3) - G(GX_RBT1) * ro) * G(GX_IT) * a;
    const double uuu = res[1] * pv[1] + res[2] * pv[2] + res[3] * pv[3] - res[4];
    const double co1 = res[0] * G(GX_AAA1) + d_ry1 * uuu + G(GX_CROSS) * res[6];
    const double co2 = res[0] * G(GX_AAA2) + d_rt * uuu + G(GX_CROSS2) * res[6];
    const double co3 = res[0] * G(GX_SSHM) + G(GX_RRR) * uuu + res[6] * G(GX_SSY2);
    const double mm = G(GX_MM);
    const double v0 = (G(GX_C4C1) * (-r6) - G(GX_C2) * co1 + G(GX_C3) * co2
                       + G(GX_RB1) * (co2 + G(GX_CROSS1) * res[5] - G(GX_CP5) * res[0])) * mm * a;
    const double v4 = (G(GX_C4IB1S1) * r6 + G(GX_C1) * co1 - G(GX_C3IB1) * co3
                       - (co3 - G(GX_S1P5) * res[0] + res[5] * G(GX_SSY1))) * mm * a;
    const double v5 = (G(GX_SS) * (res[5] - res[0] * pv[5]) - G(GX_C4SSIB1) * r6 - G(GX_C1) * co2 + G(GX_C2IB1) * co3) * mm * a;
    return bsel(n == 0, v0, bsel(n == 4, v4, bsel(n == 5, v5, bsel(n == 7, v7, bsel(n == 8, v8, v1)))));
}

HD void getx(const Par &P, int npv, const CellJ &c, const double f6[6][3], double dt,
             const double *cc, const double *tt4, const double *res, double *x)
{
    double g[NGX];
    getx_prep(P, npv, c, f6, dt, cc, tt4, g);
    for (int n = 0; n < NPV; n++) if (n < npv)
        x[n] = getx_apply([&](int q) { return g[q]; }, c.pv, c.d[D_RY1], c.d[D_RT], res, n, res[n], c.pv[n]);
}

// local pseudo time step (src/timestep_module.f90:156-198)
HD double local_dt(const Par &P, const CellJ &c, const double f6[6][3])
{
    const double *pv = c.pv, *d = c.d;
    double v2 = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
    double sp = sndp2(P, d[D_C2], v2);
    double q = sp / d[D_C2], omq = 1.0 - q;
    double eig = 0.0, msum = 0.0;
    for (int dd = 0; dd < 3; dd++) {
        double o0 = 0.5 * (f6[2 * dd][0] + f6[2 * dd + 1][0]);
        double o1 = 0.5 * (f6[2 * dd][1] + f6[2 * dd + 1][1]);
        double o2 = 0.5 * (f6[2 * dd][2] + f6[2 * dd + 1][2]);
        double s = o0 * o0 + o1 * o1 + o2 * o2;
        double uc = fabs(o0 * pv[1] + o1 * pv[2] + o2 * pv[3]);
        double e = sqrt((uc * uc) * (omq * omq) + 4.0 * sp * s);
        eig += 0.5 * (uc * (1.0 + q) + e);
        msum += s;
    }
    double ev = eigenvis(P, d, c.mut) * msum / c.vol;
    return P.cfl * c.vol / (eig + 4.0 * ev);
}

// ------------------------------------------------------------ viscous flux
// g3[q] = (normal difference, e-transverse, t-transverse) difference of
// variable vars[q] = {u, v, w, T, k, omega} over the 18-point face stencil;
// e4/t4: the four transverse face vectors of each family; gl/gr = {vol, ydns}
// of the two cells; dl/dr compact props; mutl/mutr eddy viscosities;
// cav = {uc, vc, wc, kc, oc} face averages.
template <class R>
HD void visc_flux(const Par &P, const R *nx, const R e4[4][3], const R t4[4][3], const R *gl, const R *gr,
                  const R g3[6][3], const R *cav, const R *dl, const R *dr, R mutl, R mutr, R *fx)
{
    const R c23 = R(2.0 / 3.0);
    const bool turb = (P.iturb == 0);
    R ex = R(0.25) * (e4[0][0] + e4[1][0] + e4[2][0] + e4[3][0]);
    R ey = R(0.25) * (e4[0][1] + e4[1][1] + e4[2][1] + e4[3][1]);
    R ez = R(0.25) * (e4[0][2] + e4[1][2] + e4[2][2] + e4[3][2]);
    R tx = R(0.25) * (t4[0][0] + t4[1][0] + t4[2][0] + t4[3][0]);
    R ty = R(0.25) * (t4[0][1] + t4[1][1] + t4[2][1] + t4[3][1]);
    R tz = R(0.25) * (t4[0][2] + t4[1][2] + t4[2][2] + t4[3][2]);
    R vol = R(2) / (gl[0] + gr[0]);
    R vis = R(0.5) * (dl[D_MU] + dr[D_MU]);
    R con = R(0.5) * (dl[D_KAP] + dr[D_KAP]);
    R emut = 0, tcon = 0;
    if (turb) {
        emut = R(0.5) * (mutl + mutr);
        tcon = R(0.5) * (mutl * dl[D_HT] + mutr * dr[D_HT]) / R(0.9);
    }
    R g[6][3];
    const int nv = turb ? 6 : 4;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        if (q >= nv) { g[q][0] = g[q][1] = g[q][2] = R(0); continue; }
        R a = g3[q][0], b = g3[q][1], c = g3[q][2];
        g[q][0] = (a * nx[0] + b * ex + c * tx) * vol;
        g[q][1] = (a * nx[1] + b * ey + c * ty) * vol;
        g[q][2] = (a * nx[2] + b * ez + c * tz) * vol;
    }
    R uc = cav[0], vc = cav[1], wc = cav[2];
    R mu = turb ? vis + emut : vis, kap = turb ? con + tcon : con;
    R div = g[0][0] + g[1][1] + g[2][2];
    R txx = mu * (R(2) * g[0][0] - c23 * div), tyy = mu * (R(2) * g[1][1] - c23 * div), tzz = mu * (R(2) * g[2][2] - c23 * div);
    R txy = mu * (g[0][1] + g[1][0]), txz = mu * (g[0][2] + g[2][0]), tyz = mu * (g[1][2] + g[2][1]);
    fx[0] = 0;
    fx[1] = txx * nx[0] + txy * nx[1] + txz * nx[2];
    fx[2] = txy * nx[0] + tyy * nx[1] + tyz * nx[2];
    fx[3] = txz * nx[0] + tyz * nx[1] + tzz * nx[2];
    fx[4] = nx[0] * (uc * txx + vc * txy + wc * txz + kap * g[3][0]) + nx[1] * (uc * txy + vc * tyy + wc * tyz + kap * g[3][1])
          + nx[2] * (uc * txz + vc * tyz + wc * tzz + kap * g[3][2]);
    fx[5] = 0;
    fx[6] = 0;
    if (turb) {
        R pkw = g[4][0] * g[5][0] + g[4][1] * g[5][1] + g[4][2] * g[5][2];
        R rho = R(0.5) * (dl[D_RHO] + dr[D_RHO]);
        R yd = R(0.5) * (gl[1] + gr[1]);
        R kk = cav[3], oo = cav[4];
        R cdkw = mx(R(1.712) * rho / oo * pkw, R(1.0e-10));
        R t1 = sqrt(kk) / (R(0.09) * oo * yd);
        R t2 = R(500) * vis / rho / (oo * (yd * yd));
        R t3 = (R(3.424) * rho * kk) / (cdkw * (yd * yd));
        R a1 = mn(mx(t1, t2), t3);
        R a2 = a1 * a1;
        R bf = tanh(a2 * a2);
        R tdk = bf * R(0.85) + (R(1) - bf) * R(1);
        R tdo = bf * R(0.5) + (R(1) - bf) * R(0.856);
        fx[7] = (vis + emut * tdk) * (g[4][0] * nx[0] + g[4][1] * nx[1] + g[4][2] * nx[2]);
        fx[8] = (vis + emut * tdo) * (g[5][0] * nx[0] + g[5][1] * nx[1] + g[5][2] * nx[2]);
    }
}

// ------------------------------------------------------------ cell sources
// x7[q][n]: cell (q=1), its i-, i+, j-, j+, k-, k+ neighbours (q = 0,2..6)
template <class R>
HD void gg_grad(const R x7[7][NPV], int n, const double f6[6][3], double ivol, R *g)
{
    R a1 = 0.5 * (x7[0][n] + x7[1][n]), a2 = 0.5 * (x7[2][n] + x7[1][n]);
    R a3 = 0.5 * (x7[3][n] + x7[1][n]), a4 = 0.5 * (x7[4][n] + x7[1][n]);
    R a5 = 0.5 * (x7[5][n] + x7[1][n]), a6 = 0.5 * (x7[6][n] + x7[1][n]);
    for (int m = 0; m < 3; m++)
        g[m] = (a2 * f6[1][m] + a4 * f6[3][m] + a6 * f6[5][m] - a1 * f6[0][m] - a3 * f6[2][m] - a5 * f6[4][m]) * ivol;
}

// Merkle: returns r_v - r_c (volume-integrated) and the coupling icav[4]
template <class R>
HD void merkle(const Par &P, const R *pc, const R *d, double vol, R &rc, R &rv, R *icav)
{
    rc = 0.0; rv = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = 0.0;
    if (pc[4] >= P.t_crit) return;
    R pww = sat_p(pc[4]);
    R irho = 1.0 / d[D_RHO];
    R pc_ = mx(pc[0] + P.pref - pww, R(0.0)), pv_ = mx(pww - pc[0] - P.pref, R(0.0));
    const R dpr = CavScale<R>::dp(P), tr = CavScale<R>::t(P);
    R r_c = P.c_c * d[D_RHO] * pc[5] * pc_ * dpr * tr;
    R r_v = P.c_v * d[D_RHO] * (1.0 - pc[5] - pc[6]) * pv_ * dpr * tr;
    if (P.cav_clip == 0) {   // as written in src/cav_module.f90:151-152 (pc[4] is T)
        r_c = mn(r_c, d[D_RHO] * pc[4]);
        r_v = mn(r_v, d[D_RHO] * (1.0 - pc[4] - pc[5]));
    } else {                 // on the mass fractions
        const R yl = 1.0 - pc[5] - pc[6];
        r_c = mn_rate(r_c, d[D_RHO] * pc[5], pc[5]);
        r_v = mn_rate(r_v, d[D_RHO] * yl, yl);
    }
    rc = r_c * vol;
    rv = r_v * vol;
    if (r_c != 0.0) {
        icav[0] = -(r_c * d[D_RP] * irho) - P.c_c * d[D_RHO] * pc[5] * P.dp_ref * P.t_ref;
        icav[1] = -(r_c * d[D_RT] * irho);
        icav[2] = -(r_c * d[D_RY1] * irho) - P.c_c * d[D_RHO] * pc_ * P.dp_ref * P.t_ref;
        icav[3] = -(r_c * d[D_RY2] * irho);
    }
    if (r_v != 0.0) {
        icav[0] = r_v * d[D_RP] * irho - P.c_v * d[D_RHO] * (1.0 - pc[5] - pc[6]) * P.dp_ref * P.t_ref;
        icav[1] = r_v * d[D_RT] * irho;
        icav[2] = r_v * d[D_RY1] * irho - P.c_v * d[D_RHO] * pv_ * P.dp_ref * P.t_ref;
        icav[3] = r_v * d[D_RY2] * irho - P.c_v * d[D_RHO] * pv_ * P.dp_ref * P.t_ref;
    }
    double s = (icav[2] < 0.0 || (icav[2] == 0.0 && signbit(icav[2]))) ? -1.0 : 1.0;
    for (int q = 0; q < 4; q++) icav[q] = icav[q] * s;
}

// The other mass-transfer models (src/cav_module.f90:173-357), on the full
// dv(1:18) of the cell (dv[2..4] phase densities, dv[14..17] the vapour and
// liquid density derivatives).  Rate clip and coupling sign as in merkle().
template <class R> HD void cav_clip_rates(const Par &P, const R *pc, const R *dv, R &r_c, R &r_v)
{
    if (P.cav_clip == 0) {
        r_c = mn(r_c, dv[0] * pc[4]);
        r_v = mn(r_v, dv[0] * (1.0 - pc[4] - pc[5]));
    } else {
        const R yl = 1.0 - pc[5] - pc[6];
        r_c = mn_rate(r_c, dv[0] * pc[5], pc[5]);
        r_v = mn_rate(r_v, dv[0] * yl, yl);
    }
}
template <class R> HD void cav_flip(R *icav)
{
    double s = (icav[2] < 0.0 || (icav[2] == 0.0 && signbit(icav[2]))) ? -1.0 : 1.0;
    for (int q = 0; q < 4; q++) icav[q] = icav[q] * s;
}

// Kunz (ncav 2), src/cav_module.f90:173-211
template <class R>
HD void kunz(const Par &P, const R *pc, const R *dv, double vol, R &rc, R &rv, R *icav)
{
    R pww = sat_p(pc[4]);
    R rho1 = 1.0 / dv[0];
    R av = dv[0] * pc[5] / dv[3], ag = dv[0] * pc[6] / dv[4], al = 1.0 - av - ag;
    const R dpr = CavScale<R>::dp(P), tr = CavScale<R>::t(P);
    R r_c = P.c_c * dv[3] * av * ((al + ag) * (al + ag)) * tr;
    R pv_ = mx(pww - pc[0] - P.pref, R(0.0));
    R r_v = P.c_v * dv[0] * (1.0 - pc[5] - pc[6]) * pv_ * dpr * tr;
    cav_clip_rates(P, pc, dv, r_c, r_v);
    rc = r_c * vol;
    rv = r_v * vol;
    R av1 = 1.0 - 4.0 * av + 3.0 * (av * av);
    icav[0] = -(P.c_c * P.t_ref * (dv[6] * pc[5] * av1 + 2.0 * (av * av) * dv[14] * (al + ag)));
    icav[1] = -(P.c_c * P.t_ref * (dv[7] * pc[5] * av1 + 2.0 * (av * av) * dv[15] * (al + ag)));
    icav[2] = -(P.c_c * P.t_ref * (dv[8] * pc[5] + dv[0]) * av1);
    icav[3] = -(P.c_c * P.t_ref * dv[9] * pc[5] * av1);
    if (r_v != 0.0) {
        icav[0] = r_v * dv[6] * rho1 - P.c_v * dv[0] * (1.0 - pc[5] - pc[6]) * P.dp_ref * P.t_ref;
        icav[1] = r_v * dv[7] * rho1;
        icav[2] = r_v * dv[8] * rho1 - P.c_v * dv[0] * pv_ * P.dp_ref * P.t_ref;
        icav[3] = r_v * dv[9] * rho1 - P.c_v * dv[0] * pv_ * P.dp_ref * P.t_ref;
    }
}

// Singhal (ncav 3), src/cav_module.f90:213-260
template <class R>
HD void singhal(const Par &P, const R *pc, const R *dv, double vol, R &rc, R &rv, R *icav)
{
    R pww = sat_p(pc[4]);
    R sigma = h2o_sigma(pc[4]);
    R rho1 = 1.0 / dv[2], rho2 = 1.0 / dv[3];
    const R uref = RefSpeed<R>::of(P);      // cav%uref is the config uref (src/cav_module.f90:77, 228)
    R kc = P.c_c * uref / sigma * dv[2] * dv[3], kv = P.c_v * uref / sigma * dv[2] * dv[3];
    R pc_ = mx(pc[0] + P.pref - pww, R(0.0)), pv_ = mx(pww - pc[0] - P.pref, R(0.0));
    R r_c = kc * sqrt(2.0 / 3.0 * pc_ * rho1) * pc[5];
    R r_v = kv * sqrt(2.0 / 3.0 * pv_ * rho1) * (1.0 - pc[5] - pc[6]);
    cav_clip_rates(P, pc, dv, r_c, r_v);
    rc = r_c * vol;
    rv = r_v * vol;
    if (r_c != 0.0) {
        icav[0] = -(r_c * (dv[14] * rho2 + 0.5 * dv[16] * rho1))
                  - 0.5 * P.c_c * P.uref / sigma * dv[2] * dv[3] * sqrt(2.0 / 3.0 * rho1) * pc[5] * pow(pc[0] + P.pref - pww, -0.5);
        icav[1] = -(r_c * (dv[15] * rho2 + 0.5 * dv[17] * rho1));
        icav[2] = -(kc * sqrt(2.0 / 3.0 * pc_ * rho1));
        icav[3] = R(0.0);
    }
    if (r_v != 0.0) {
        icav[0] = r_v * (dv[14] * rho2 + 0.5 * dv[16] * rho1)
                  - 0.5 * P.c_v * P.uref / sigma * dv[2] * dv[3] * sqrt(2.0 / 3.0 * rho1) * (1.0 - pc[5] - pc[6])
                        * pow(pww - pc[0] - P.pref, -0.5);
        icav[1] = r_v * (dv[15] * rho2 + 0.5 * dv[17] * rho1);
        icav[2] = -(kv * sqrt(2.0 / 3.0 * pv_ * rho1));
        icav[3] = -(kv * sqrt(2.0 / 3.0 * pv_ * rho1));
    }
}

// Schnerr-Sauer (ncav 4), src/cav_module.f90:262-316; bubble number density 1e13
template <class R>
HD void schnerr_sauer(const Par &P, const R *pc, const R *dv, double vol, R &rc, R &rv, R *icav)
{
    const double pi = 4.0 * atan(1.0), eta = 1.0e13;
    R pww = sat_p(pc[4]);
    R rho1 = 1.0 / dv[0], rho3 = 1.0 / dv[2], rho4 = 1.0 / dv[3];
    R y1 = pc[5], y2 = pc[6], yl = 1.0 - pc[5] - pc[6];
    double kb = pow(4.0 * pi * eta / 3.0, 1.0 / 3.0), k23 = sqrt(2.0 / 3.0);
    R r_c = P.c_c * 3.0 * kb * k23 * dv[0] * pow(dv[3], 1.0 / 3.0) * pow(dv[2], -5.0 / 6.0)
                 * pow(y1, 2.0 / 3.0) * pow(yl, 4.0 / 3.0) * sqrt(mx(pc[0] + P.pref - pww, R(0.0)));
    R r_v = P.c_v * 3.0 * kb * k23 * dv[0] * pow(dv[3], 1.0 / 3.0) * pow(dv[2], -5.0 / 6.0)
                 * pow(y1 + y2, 2.0 / 3.0) * pow(yl, 4.0 / 3.0) * sqrt(mx(pww - pc[0] - P.pref, R(0.0)));
    if constexpr (IsDual<R>::v) {
        // at y1 = 0 on the condensing side r_c grows as y1^(2/3) (infinite
        // slope, which the dual pow reports as 0), so for y1 > 0 the species
        // clip rho*y1 is the smaller: take its slope (same value, 0)
        if (P.cav_clip == 1 && y1 == 0.0 && pc[0] + P.pref - pww > 0.0 && yl > 0.0) r_c = dv[0] * pc[5];
    }
    cav_clip_rates(P, pc, dv, r_c, r_v);
    rc = r_c * vol;
    rv = r_v * vol;
    if (r_c != 0.0) {
        icav[0] = -(r_c * (dv[6] * rho1 + dv[14] * rho4 / 3.0 - 5.0 * dv[16] * rho3 / 6.0 + 0.5 / (pc[0] + P.pref - pww)));
        icav[1] = -(r_c * (dv[7] * rho1 + dv[15] * rho4 / 3.0 - 5.0 * dv[17] * rho3 / 6.0));
        icav[2] = -(r_c * (dv[8] * rho1 + 2.0 / (3.0 * y1) - 4.0 / (3.0 * yl)));
        icav[3] = -(r_c * (dv[9] * rho1 - 4.0 / (3.0 * yl)));
    }
    if (r_v != 0.0) {
        icav[0] = r_v * (dv[6] * rho1 + dv[14] * rho4 / 3.0 - 5.0 * dv[16] * rho3 / 6.0 - 0.5 / (pww - pc[0] - P.pref));
        icav[1] = r_v * (dv[7] * rho1 + dv[15] * rho4 / 3.0 - 5.0 * dv[17] * rho3 / 6.0);
        icav[2] = r_v * (dv[8] * rho1 + 2.0 / (3.0 * (y1 + y2)) - 4.0 / (3.0 * yl));
        icav[3] = r_v * (dv[9] * rho1 + 2.0 / (3.0 * (y1 + y2)) - 4.0 / (3.0 * yl));
    }
}

// Lee (ncav 5), src/cav_module.f90:318-356
template <class R>
HD void lee(const Par &P, const R *pc, const R *dv, double vol, R &rc, R &rv, R *icav)
{
    R tww = sat_t(pc[0] + P.pref);
    R rho1 = 1.0 / dv[0];
    R tc = mx(tww - pc[4], R(0.0)), tv_ = mx(pc[4] - tww, R(0.0));
    R r_c = P.c_c * dv[0] * pc[5] * tc / tww;
    R r_v = P.c_v * dv[0] * (1.0 - pc[5] - pc[6]) * tv_ / tww;
    cav_clip_rates(P, pc, dv, r_c, r_v);
    rc = r_c * vol;
    rv = r_v * vol;
    if (r_c != 0.0) {
        icav[0] = -(r_c * dv[6] * rho1);
        icav[1] = -(r_c * dv[7] * rho1) + P.c_c * dv[0] * pc[5] / tww;
        icav[2] = -(r_c * dv[8] * rho1) - P.c_c * dv[0] * tc / tww;
        icav[3] = -(r_c * dv[9] * rho1);
    }
    if (r_v != 0.0) {
        icav[0] = r_v * dv[6] * rho1;
        icav[1] = r_v * dv[7] * rho1 + P.c_v * dv[0] * (1.0 - pc[5] - pc[6]) / tww;
        icav[2] = r_v * dv[8] * rho1 - P.c_v * dv[0] * tv_ / tww;
        icav[3] = r_v * dv[9] * rho1 - P.c_v * dv[0] * tv_ / tww;
    }
}

// the model of ncav 2..5 (src/rhs_module.f90:98-110), full dv(1:18) of the cell
template <class R>
HD void cav_model(const Par &P, const R *pc, const R *dv, double vol, R &rc, R &rv, R *icav)
{
    rc = 0.0; rv = 0.0;
    icav[0] = icav[1] = icav[2] = icav[3] = R(0.0);
    if (pc[4] >= P.t_crit) return;
    if (P.ncav == 2) kunz(P, pc, dv, vol, rc, rv, icav);
    else if (P.ncav == 3) singhal(P, pc, dv, vol, rc, rv, icav);
    else if (P.ncav == 4) schnerr_sauer(P, pc, dv, vol, rc, rv, icav);
    else lee(P, pc, dv, vol, rc, rv, icav);
    cav_flip(icav);
}

// Turbulent length scale of the SST destruction term (calturblength,
// src/turbsource_module.f90:406-576): RANS sqrt(k)/(9 omega)*100, or its DES
// (1), DDES (2), IDDES (3) or S-IDDES (4) blend with the LES scale C_DES h_max,
// h from the cell volume over its mean face areas.  Constant integer powers
// as gfortran expands them (x**3 = (x*x)*x, x**10 = x5*x5 with x5 = x3*x2).
template <class R>
HD R des_length(int ides, R k, R om, const double f6[6][3], double vol, double ydns, R mu, R mut, R bigf, R s2, R w2)
{
    R tl_rans = sqrt(k) / 9.0 / om * 100.0;
    if (ides == 0) return tl_rans;
    double gs[3];
    for (int q = 0; q < 3; q++) {
        const double *a = f6[2 * q], *b = f6[2 * q + 1];
        gs[q] = 0.5 * (sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) + sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
        gs[q] = vol / gs[q];
    }
    double hmax = mx(mx(gs[0], gs[1]), gs[2]);
    if (ides == 1) return mn(tl_rans, R(0.61 * hmax));
    R cdes = 0.78 * bigf + 0.61 * (1.0 - bigf);
    R den = 0.41 * 0.41 * (ydns * ydns) * sqrt(0.5 * (s2 + w2));
    if (ides == 2) {
        R tl_les = cdes * hmax;
        R rd = (mu + mut) / den, x = 20.0 * rd;
        R fd = 1.0 - tanh(x * x * x);
        return tl_rans - fd * mx(R(0.0), tl_rans - tl_les);
    }
    R tl_les = cdes * mn(0.15 * mx(ydns, hmax), hmax);
    double alpha = 0.25 - ydns / hmax;
    double fb = mn(2.0 * exp(-9.0 * (alpha * alpha)), 1.0);
    R rdt = mut / den, xt = 20.0 * rdt;
    R fdt = 1.0 - tanh(xt * xt * xt);
    R fd = mx(1.0 - fdt, R(fb));
    if (ides == 4) return fd * tl_rans + (1.0 - fd) * tl_les;
    R rdl = mu / den;
    R xl = 25.0 * rdl, xl2 = xl * xl, xl3 = xl2 * xl, xl5 = xl3 * xl2;
    R fl = tanh(xl5 * xl5);
    R x3 = 3.4969 * rdt;
    R ft = tanh(x3 * x3 * x3);
    R fe2 = 1.0 - mx(ft, fl);
    double fe1 = alpha >= 0.0 ? 2.0 * exp(-11.09 * (alpha * alpha)) : 2.0 * exp(-9.0 * (alpha * alpha));
    R fe = fe2 * mx(fe1 - 1.0, 0.0);
    return fd * (1.0 + fe) * tl_rans + (1.0 - fd) * tl_les;
}

// The SST source from the cell's Green-Gauss gradients of u, v, w, k, omega
// and its own k, omega (the adjoint differentiates this form: the neighbours
// enter only through the gradients, which are linear in them).
template <bool DES = false, class R = double>
HD void sst_source_g(const Par &P, const R *gu, const R *gv, const R *gw, const R *gk, const R *go, R k, R om,
                     const double f6[6][3], double vol, double ydns, const R *d, R mut, R *src, R *itt, R &ocut)
{
    const double c23 = 2.0 / 3.0, c59 = 5.0 / 9.0;
    R div = gu[0] + gv[1] + gw[2];
    R s2 = 2.0 * (gu[0] * gu[0] + gv[1] * gv[1] + gw[2] * gw[2]) + (gu[1] + gv[0]) * (gu[1] + gv[0])
         + (gu[2] + gw[0]) * (gu[2] + gw[0]) + (gv[2] + gw[1]) * (gv[2] + gw[1]);
    R rho = d[D_RHO];
    R prod = mut * (s2 - c23 * (div * div)) - c23 * rho * k * div;
    prod = mn(prod, 0.9 * rho * k * om);
    ocut = sqrt(s2 - c23 * (div * div));
    R pkw = gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2];
    R cdkw = mx(1.712 * rho / om * pkw, R(1.0e-10));
    R t1 = sqrt(k) / (0.09 * om * ydns);
    R t2 = 500.0 * d[D_MU] / rho / (om * (ydns * ydns));
    R t3 = (3.424 * rho * k) / (cdkw * (ydns * ydns));
    R a1 = mn(mx(t1, t2), t3);
    R a2 = a1 * a1;
    R bf = tanh(a2 * a2);
    R ta = bf * c59 + (1.0 - bf) * 0.44;
    R tb = bf * 0.075 + (1.0 - bf) * 0.0828;
    R ki = 1.0 / k, oi = 1.0 / om, ei = 1.0 / mut;
    R tl;
    if constexpr (DES) {
        R w2 = (gu[1] - gv[0]) * (gu[1] - gv[0]) + (gu[2] - gw[0]) * (gu[2] - gw[0]) + (gv[2] - gw[1]) * (gv[2] - gw[1]);
        tl = des_length<R>(P.ides, k, om, f6, vol, ydns, d[D_MU], mut, bf, s2, w2);
    } else {
        tl = sqrt(k) / 9.0 / om * 100.0;
    }
    R dest = rho * pow(k, 1.5) / tl;
    src[0] = (prod - dest) * vol;
    src[1] = (ta * rho * ei * prod - tb * rho * (om * om) + 1.712 * (1.0 - bf) * rho * oi * pkw) * vol;
    R dpdk = prod * ki;
    R dpdo = -((prod + c23 * rho * k * div) * oi);
    itt[0] = dpdk - dest * ki;
    itt[1] = dpdo - dest * oi;
    itt[2] = 0.0;
    itt[3] = ta * rho * ei * dpdo + ta * rho * ei * prod * oi - 2.0 * tb * rho * om - 1.712 * (1.0 - bf) * rho * (oi * oi) * pkw;
    R t = itt[0] + itt[3], det = itt[0] * itt[3] - itt[1] * itt[2];
    R sq = sqrt(mx(0.25 * (t * t) - det, R(0.0)));
    R l1 = 0.5 * t + sq, l2 = 0.5 * t - sq;
    double s1 = signbit(l1) ? -1.0 : 1.0, s2s = signbit(l2) ? -1.0 : 1.0;
    itt[0] = itt[0] * s1; itt[1] = itt[1] * s1; itt[2] = itt[2] * s2s; itt[3] = itt[3] * s2s;
}

// k-omega SST source: src[2], itt[4], omega_cut.  DES: the length-scale
// variants are compiled in (the RANS-only instantiation keeps its registers).
template <bool DES = false, class R = double>
HD void sst_source(const Par &P, const R x7[7][NPV], const double f6[6][3], double vol, double ydns, const R *d,
                   R mut, R *src, R *itt, R &ocut)
{
    double iv = 1.0 / vol;
    R gu[3], gv[3], gw[3], gk[3], go[3];
    gg_grad(x7, 7, f6, iv, gk);
    gg_grad(x7, 8, f6, iv, go);
    gg_grad(x7, 1, f6, iv, gu);
    gg_grad(x7, 2, f6, iv, gv);
    gg_grad(x7, 3, f6, iv, gw);
    sst_source_g<DES, R>(P, gu, gv, gw, gk, go, x7[1][7], x7[1][8], f6, vol, ydns, d, mut, src, itt, ocut);
}

// SST eddy viscosity mu_t = 0.31 rho k / max(0.31 omega, S F2)
// the SST eddy viscosity from the cell's rho, mu, k, omega and its velocity
// gradients gu, gv, gw (src/eddy_module.f90:135-187); R = double or a dual
// number (the adjoint differentiates it with respect to all eleven inputs)
template <class R>
HD R sst_eddy_g(R rho, R mu, R k, R om, const R *gu, const R *gv, const R *gw, double ydns)
{
    R t1 = sqrt(k) / (0.09 * om * ydns);
    R t2 = 500.0 * mu / rho / (om * (ydns * ydns));
    R a2 = mx(2.0 * t1, t2);
    R bf2 = tanh(a2 * a2);
    R sij = gu[0] * gu[0] + gv[1] * gv[1] + gw[2] * gw[2]
          + 0.5 * ((gu[1] + gv[0]) * (gu[1] + gv[0]) + (gu[2] + gw[0]) * (gu[2] + gw[0]) + (gv[2] + gw[1]) * (gv[2] + gw[1]));
    R sss = sqrt(2.0 * sij);
    return 0.31 * rho * k / mx(0.31 * om, sss * bf2);
}

HD double sst_eddy(const double x7[7][NPV], const double f6[6][3], double vol, double ydns, const double *d)
{
    double iv = 1.0 / vol;
    double gu[3], gv[3], gw[3];
    gg_grad(x7, 1, f6, iv, gu);
    gg_grad(x7, 2, f6, iv, gv);
    gg_grad(x7, 3, f6, iv, gw);
    return sst_eddy_g<double>(d[D_RHO], d[D_MU], x7[1][7], x7[1][8], gu, gv, gw, ydns);
}

}  // namespace cav
