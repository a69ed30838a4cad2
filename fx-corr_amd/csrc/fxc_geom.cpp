// Station geometry on the host: the source's RA/Dec/epoch parsing, J2000 ->
// mean-of-date precession, Greenwich sidereal time and the baseline
// geometric delay with its rate and acceleration (geom.rs), and the
// 1-second geometric-delay table that feeds the per-frame delay model
// (main.rs:1628-1642).
//
// The reference calls ERFA (erfa 0.2.1): eraPmat06 (IAU 2006 precession with
// frame bias), eraGst06a (apparent sidereal time, IAU 2006/2000A), eraAnp,
// eraS2c, eraC2s, eraRxp.  Those routines are restated here from the
// published models:
//   - Earth rotation angle, IAU 2000 (Capitaine et al. 2000);
//   - GMST, IAU 2006 (Capitaine, Wallace & Chapront 2005), the polynomial in
//     TT added to the rotation angle;
//   - precession, IAU 2006 Fukushima-Williams angles (Hilton et al. 2006)
//     and mean obliquity of date;
//   - apparent sidereal time as GMST plus the equation of the equinoxes,
//     dpsi cos(eps_A) plus the complementary terms (IERS Conventions 2003,
//     ch. 5), with dpsi from the IAU 2000B series (McCarthy & Luzum 2003:
//     77 luni-solar terms and a fixed planetary offset) scaled by the IAU
//     2006 adjustments (Wallace & Capitaine 2006), in place of the 1365-term
//     IAU 2000A series.  2000B is within 1 mas of 2000A; at 2006-01-01 the
//     sidereal time differs from ERFA's eraGst06a by 1.5e-9 rad, i.e.
//     5e-18 s of delay per metre of baseline.
// Everything after the sidereal time (source vector, delays, central
// differences) follows geom.rs operation by operation, without contraction.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fxcorr.h"

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);

#pragma clang fp contract(off)

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double k2Pi = 6.283185307179586476925287;
constexpr double kAs2R = 4.848136811095359935899141e-6;   // arcseconds to radians
constexpr double kTurnAs = 1296000.0;                      // arcseconds in a full circle
constexpr double kJ2000 = 2451545.0;                       // JD of J2000.0
constexpr double kDaysPerCentury = 36525.0;
constexpr double kMjdZero = 2400000.5;
constexpr double kC = 299792458.0;                         // geom.rs:4

double anp(double a) {                                      // angle into [0, 2 pi)
    double w = std::fmod(a, k2Pi);
    if (w < 0) w += k2Pi;
    return w;
}

// Rust's f64::to_radians: x * (pi / 180) with the quotient rounded once.
double to_radians(double deg) { return deg * (kPi / 180.0); }

// ------------------------------------------------------------ time scales
// Earth rotation angle (IAU 2000) for a two-part Julian date.
double earth_rotation_angle(double dj1, double dj2) {
    const double d1 = dj1 < dj2 ? dj1 : dj2, d2 = dj1 < dj2 ? dj2 : dj1;
    const double t = d1 + (d2 - kJ2000);
    const double f = std::fmod(d1, 1.0) + std::fmod(d2, 1.0);
    return anp(k2Pi * (f + 0.7790572732640 + 0.00273781191135448 * t));
}

double centuries_tt(double d1, double d2) { return ((d1 - kJ2000) + d2) / kDaysPerCentury; }

// GMST (IAU 2006): rotation angle plus a polynomial in TT centuries.
double gmst_2006(double ut1, double ut2, double tt1, double tt2) {
    const double t = centuries_tt(tt1, tt2);
    const double poly = 0.014506 +
                        (4612.156534 + (1.3915817 + (-0.00000044 + (-0.000029956 + (-0.0000000368) * t) * t) * t) * t) * t;
    return anp(earth_rotation_angle(ut1, ut2) + poly * kAs2R);
}

// Mean obliquity of the ecliptic (IAU 2006).
double mean_obliquity_2006(double t) {
    return (84381.406 + (-46.836769 + (-0.0001831 + (0.00200340 + (-0.000000576 + (-0.0000000434) * t) * t) * t) * t) * t) *
           kAs2R;
}

// ------------------------------------------------------------ precession
using Mat3 = double[3][3];

void rot_z(double a, Mat3 r) {   // r = Rz(a) r
    const double s = std::sin(a), c = std::cos(a);
    for (int j = 0; j < 3; ++j) {
        const double x = r[0][j], y = r[1][j];
        r[0][j] = c * x + s * y;
        r[1][j] = -s * x + c * y;
    }
}

void rot_x(double a, Mat3 r) {   // r = Rx(a) r
    const double s = std::sin(a), c = std::cos(a);
    for (int j = 0; j < 3; ++j) {
        const double y = r[1][j], z = r[2][j];
        r[1][j] = c * y + s * z;
        r[2][j] = -s * y + c * z;
    }
}

// Bias-precession matrix from the IAU 2006 Fukushima-Williams angles:
// R = Rx(-eps_A) Rz(-psi) Rx(phi) Rz(gamma).
void bias_precession_2006(double d1, double d2, Mat3 r) {
    const double t = centuries_tt(d1, d2);
    const double gamb =
        (-0.052928 + (10.556378 + (0.4932044 + (-0.00031238 + (-0.000002788 + (0.0000000260) * t) * t) * t) * t) * t) * kAs2R;
    const double phib =
        (84381.412819 + (-46.811016 + (0.0511268 + (0.00053289 + (-0.000000440 + (-0.0000000176) * t) * t) * t) * t) * t) *
        kAs2R;
    const double psib =
        (-0.041775 + (5038.481484 + (1.5584175 + (-0.00018522 + (-0.000026452 + (-0.0000000148) * t) * t) * t) * t) * t) * kAs2R;
    const double epsa = mean_obliquity_2006(t);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = (i == j) ? 1.0 : 0.0;
    rot_z(gamb, r);
    rot_x(phib, r);
    rot_z(-psib, r);
    rot_x(-epsa, r);
}

// ------------------------------------------------------------ nutation
// Delaunay arguments (IERS Conventions 2003), radians.
struct Delaunay { double l, lp, f, d, om; };

Delaunay delaunay(double t) {
    Delaunay a;
    a.l = std::fmod(485868.249036 + t * (1717915923.2178 + t * (31.8792 + t * (0.051635 + t * (-0.00024470)))), kTurnAs) * kAs2R;
    a.lp = std::fmod(1287104.793048 + t * (129596581.0481 + t * (-0.5532 + t * (0.000136 + t * (-0.00001149)))), kTurnAs) * kAs2R;
    a.f = std::fmod(335779.526232 + t * (1739527262.8478 + t * (-12.7512 + t * (-0.001037 + t * (0.00000417)))), kTurnAs) * kAs2R;
    a.d = std::fmod(1072260.703692 + t * (1602961601.2090 + t * (-6.3706 + t * (0.006593 + t * (-0.00003169)))), kTurnAs) * kAs2R;
    a.om = std::fmod(450160.398036 + t * (-6962890.5431 + t * (7.4722 + t * (0.007702 + t * (-0.00005939)))), kTurnAs) * kAs2R;
    return a;
}

// IAU 2000B luni-solar nutation (McCarthy & Luzum 2003): multipliers of
// l, l', F, D, Om; longitude sin, sin*t, cos; obliquity cos, cos*t, sin
// (units of 0.1 microarcsecond).
struct NutTerm { signed char n[5]; double ps, pst, pc, ec, ect, es; };
const NutTerm kNut2000B[77] = {
    {{0, 0, 0, 0, 1}, -172064161.0, -174666.0, 33386.0, 92052331.0, 9086.0, 15377.0},
    {{0, 0, 2, -2, 2}, -13170906.0, -1675.0, -13696.0, 5730336.0, -3015.0, -4587.0},
    {{0, 0, 2, 0, 2}, -2276413.0, -234.0, 2796.0, 978459.0, -485.0, 1374.0},
    {{0, 0, 0, 0, 2}, 2074554.0, 207.0, -698.0, -897492.0, 470.0, -291.0},
    {{0, 1, 0, 0, 0}, 1475877.0, -3633.0, 11817.0, 73871.0, -184.0, -1924.0},
    {{0, 1, 2, -2, 2}, -516821.0, 1226.0, -524.0, 224386.0, -677.0, -174.0},
    {{1, 0, 0, 0, 0}, 711159.0, 73.0, -872.0, -6750.0, 0.0, 358.0},
    {{0, 0, 2, 0, 1}, -387298.0, -367.0, 380.0, 200728.0, 18.0, 318.0},
    {{1, 0, 2, 0, 2}, -301461.0, -36.0, 816.0, 129025.0, -63.0, 367.0},
    {{0, -1, 2, -2, 2}, 215829.0, -494.0, 111.0, -95929.0, 299.0, 132.0},
    {{0, 0, 2, -2, 1}, 128227.0, 137.0, 181.0, -68982.0, -9.0, 39.0},
    {{-1, 0, 2, 0, 2}, 123457.0, 11.0, 19.0, -53311.0, 32.0, -4.0},
    {{-1, 0, 0, 2, 0}, 156994.0, 10.0, -168.0, -1235.0, 0.0, 82.0},
    {{1, 0, 0, 0, 1}, 63110.0, 63.0, 27.0, -33228.0, 0.0, -9.0},
    {{-1, 0, 0, 0, 1}, -57976.0, -63.0, -189.0, 31429.0, 0.0, -75.0},
    {{-1, 0, 2, 2, 2}, -59641.0, -11.0, 149.0, 25543.0, -11.0, 66.0},
    {{1, 0, 2, 0, 1}, -51613.0, -42.0, 129.0, 26366.0, 0.0, 78.0},
    {{-2, 0, 2, 0, 1}, 45893.0, 50.0, 31.0, -24236.0, -10.0, 20.0},
    {{0, 0, 0, 2, 0}, 63384.0, 11.0, -150.0, -1220.0, 0.0, 29.0},
    {{0, 0, 2, 2, 2}, -38571.0, -1.0, 158.0, 16452.0, -11.0, 68.0},
    {{0, -2, 2, -2, 2}, 32481.0, 0.0, 0.0, -13870.0, 0.0, 0.0},
    {{-2, 0, 0, 2, 0}, -47722.0, 0.0, -18.0, 477.0, 0.0, -25.0},
    {{2, 0, 2, 0, 2}, -31046.0, -1.0, 131.0, 13238.0, -11.0, 59.0},
    {{1, 0, 2, -2, 2}, 28593.0, 0.0, -1.0, -12338.0, 10.0, -3.0},
    {{-1, 0, 2, 0, 1}, 20441.0, 21.0, 10.0, -10758.0, 0.0, -3.0},
    {{2, 0, 0, 0, 0}, 29243.0, 0.0, -74.0, -609.0, 0.0, 13.0},
    {{0, 0, 2, 0, 0}, 25887.0, 0.0, -66.0, -550.0, 0.0, 11.0},
    {{0, 1, 0, 0, 1}, -14053.0, -25.0, 79.0, 8551.0, -2.0, -45.0},
    {{-1, 0, 0, 2, 1}, 15164.0, 10.0, 11.0, -8001.0, 0.0, -1.0},
    {{0, 2, 2, -2, 2}, -15794.0, 72.0, -16.0, 6850.0, -42.0, -5.0},
    {{0, 0, -2, 2, 0}, 21783.0, 0.0, 13.0, -167.0, 0.0, 13.0},
    {{1, 0, 0, -2, 1}, -12873.0, -10.0, -37.0, 6953.0, 0.0, -14.0},
    {{0, -1, 0, 0, 1}, -12654.0, 11.0, 63.0, 6415.0, 0.0, 26.0},
    {{-1, 0, 2, 2, 1}, -10204.0, 0.0, 25.0, 5222.0, 0.0, 15.0},
    {{0, 2, 0, 0, 0}, 16707.0, -85.0, -10.0, 168.0, -1.0, 10.0},
    {{1, 0, 2, 2, 2}, -7691.0, 0.0, 44.0, 3268.0, 0.0, 19.0},
    {{-2, 0, 2, 0, 0}, -11024.0, 0.0, -14.0, 104.0, 0.0, 2.0},
    {{0, 1, 2, 0, 2}, 7566.0, -21.0, -11.0, -3250.0, 0.0, -5.0},
    {{0, 0, 2, 2, 1}, -6637.0, -11.0, 25.0, 3353.0, 0.0, 14.0},
    {{0, -1, 2, 0, 2}, -7141.0, 21.0, 8.0, 3070.0, 0.0, 4.0},
    {{0, 0, 0, 2, 1}, -6302.0, -11.0, 2.0, 3272.0, 0.0, 4.0},
    {{1, 0, 2, -2, 1}, 5800.0, 10.0, 2.0, -3045.0, 0.0, -1.0},
    {{2, 0, 2, -2, 2}, 6443.0, 0.0, -7.0, -2768.0, 0.0, -4.0},
    {{-2, 0, 0, 2, 1}, -5774.0, -11.0, -15.0, 3041.0, 0.0, -5.0},
    {{2, 0, 2, 0, 1}, -5350.0, 0.0, 21.0, 2695.0, 0.0, 12.0},
    {{0, -1, 2, -2, 1}, -4752.0, -11.0, -3.0, 2719.0, 0.0, -3.0},
    {{0, 0, 0, -2, 1}, -4940.0, -11.0, -21.0, 2720.0, 0.0, -9.0},
    {{-1, -1, 0, 2, 0}, 7350.0, 0.0, -8.0, -51.0, 0.0, 4.0},
    {{2, 0, 0, -2, 1}, 4065.0, 0.0, 6.0, -2206.0, 0.0, 1.0},
    {{1, 0, 0, 2, 0}, 6579.0, 0.0, -24.0, -199.0, 0.0, 2.0},
    {{0, 1, 2, -2, 1}, 3579.0, 0.0, 5.0, -1900.0, 0.0, 1.0},
    {{1, -1, 0, 0, 0}, 4725.0, 0.0, -6.0, -41.0, 0.0, 3.0},
    {{-2, 0, 2, 0, 2}, -3075.0, 0.0, -2.0, 1313.0, 0.0, -1.0},
    {{3, 0, 2, 0, 2}, -2904.0, 0.0, 15.0, 1233.0, 0.0, 7.0},
    {{0, -1, 0, 2, 0}, 4348.0, 0.0, -10.0, -81.0, 0.0, 2.0},
    {{1, -1, 2, 0, 2}, -2878.0, 0.0, 8.0, 1232.0, 0.0, 4.0},
    {{0, 0, 0, 1, 0}, -4230.0, 0.0, 5.0, -20.0, 0.0, -2.0},
    {{-1, -1, 2, 2, 2}, -2819.0, 0.0, 7.0, 1207.0, 0.0, 3.0},
    {{-1, 0, 2, 0, 0}, -4056.0, 0.0, 5.0, 40.0, 0.0, -2.0},
    {{0, -1, 2, 2, 2}, -2647.0, 0.0, 11.0, 1129.0, 0.0, 5.0},
    {{-2, 0, 0, 0, 1}, -2294.0, 0.0, -10.0, 1266.0, 0.0, -4.0},
    {{1, 1, 2, 0, 2}, 2481.0, 0.0, -7.0, -1062.0, 0.0, -3.0},
    {{2, 0, 0, 0, 1}, 2179.0, 0.0, -2.0, -1129.0, 0.0, -2.0},
    {{-1, 1, 0, 1, 0}, 3276.0, 0.0, 1.0, -9.0, 0.0, 0.0},
    {{1, 1, 0, 0, 0}, -3389.0, 0.0, 5.0, 35.0, 0.0, -2.0},
    {{1, 0, 2, 0, 0}, 3339.0, 0.0, -13.0, -107.0, 0.0, 1.0},
    {{-1, 0, 2, -2, 1}, -1987.0, 0.0, -6.0, 1073.0, 0.0, -2.0},
    {{1, 0, 0, 0, 2}, -1981.0, 0.0, 0.0, 854.0, 0.0, 0.0},
    {{-1, 0, 0, 1, 0}, 4026.0, 0.0, -353.0, -553.0, 0.0, -139.0},
    {{0, 0, 2, 1, 2}, 1660.0, 0.0, -5.0, -710.0, 0.0, -2.0},
    {{-1, 0, 2, 4, 2}, -1521.0, 0.0, 9.0, 647.0, 0.0, 4.0},
    {{-1, 1, 0, 1, 1}, 1314.0, 0.0, 0.0, -700.0, 0.0, 0.0},
    {{0, -2, 2, -2, 1}, -1283.0, 0.0, 0.0, 672.0, 0.0, 0.0},
    {{1, 2, 2, -2, 2}, -1331.0, 0.0, 8.0, 663.0, 0.0, 4.0},
    {{-2, 0, 2, 2, 2}, 1383.0, 0.0, -2.0, -594.0, 0.0, -2.0},
    {{-1, 0, 0, 0, 2}, 1405.0, 0.0, 4.0, -610.0, 0.0, 2.0},
    {{1, 1, 2, -2, 2}, 1290.0, 0.0, 0.0, -556.0, 0.0, 0.0},
};

// Nutation in longitude (IAU 2000B), radians.  The series evaluates its
// arguments with the linear Delaunay rates, as the 2000B model specifies.
double nutation_longitude_2000b(double t) {
    const double el = std::fmod(485868.249036 + 1717915923.2178 * t, kTurnAs) * kAs2R;
    const double elp = std::fmod(1287104.79305 + 129596581.0481 * t, kTurnAs) * kAs2R;
    const double f = std::fmod(335779.526232 + 1739527262.8478 * t, kTurnAs) * kAs2R;
    const double d = std::fmod(1072260.70369 + 1602961601.2090 * t, kTurnAs) * kAs2R;
    const double om = std::fmod(450160.398036 + -6962890.5431 * t, kTurnAs) * kAs2R;
    double dp = 0.0;
    for (int i = 76; i >= 0; --i) {   // smallest terms first
        const NutTerm &x = kNut2000B[i];
        const double arg = std::fmod((double)x.n[0] * el + (double)x.n[1] * elp + (double)x.n[2] * f +
                                         (double)x.n[3] * d + (double)x.n[4] * om,
                                     k2Pi);
        dp += (x.ps + x.pst * t) * std::sin(arg) + x.pc * std::cos(arg);
    }
    const double planetary = -0.135 * (kAs2R / 1.0e3);   // fixed offset standing in for the planetary terms
    return dp * (kAs2R / 1.0e7) + planetary;
}

// Complementary terms of the equation of the equinoxes (IERS 2003), the
// terms above 0.5 microarcsecond, radians.
double equinox_complement(double t) {
    struct EqTerm { signed char n[5]; double s, c; };
    static const EqTerm e0[12] = {
        {{0, 0, 0, 0, 1}, 2640.96e-6, -0.39e-6}, {{0, 0, 0, 0, 2}, 63.52e-6, -0.02e-6},
        {{0, 0, 2, -2, 3}, 11.75e-6, 0.01e-6},   {{0, 0, 2, -2, 1}, 11.21e-6, 0.01e-6},
        {{0, 0, 2, -2, 2}, -4.55e-6, 0.0},       {{0, 0, 2, 0, 3}, 2.02e-6, 0.0},
        {{0, 0, 2, 0, 1}, 1.98e-6, 0.0},         {{0, 0, 0, 0, 3}, -1.72e-6, 0.0},
        {{0, 1, 0, 0, 1}, -1.41e-6, -0.01e-6},   {{0, 1, 0, 0, -1}, -1.26e-6, -0.01e-6},
        {{1, 0, 0, 0, -1}, -0.63e-6, 0.0},       {{1, 0, 0, 0, 1}, -0.63e-6, 0.0},
    };
    const Delaunay a = delaunay(t);
    const double fa[5] = {a.l, a.lp, a.f, a.d, a.om};
    double s0 = 0.0;
    for (int i = 11; i >= 0; --i) {
        double arg = 0.0;
        for (int j = 0; j < 5; ++j) arg += (double)e0[i].n[j] * fa[j];
        s0 += e0[i].s * std::sin(arg) + e0[i].c * std::cos(arg);
    }
    const double s1 = -0.87e-6 * std::sin(a.om);
    return (s0 + s1 * t) * kAs2R;
}

// Apparent sidereal time: GMST (IAU 2006) + dpsi cos(eps_A) + complement,
// dpsi the 2000B series with the IAU 2006 adjustments (J2 secular rate and
// the 0.4697 ppm scale).
double gast_2006(double ut1, double ut2, double tt1, double tt2) {
    const double t = centuries_tt(tt1, tt2);
    const double dp00 = nutation_longitude_2000b(t);
    const double dpsi = dp00 + dp00 * (0.4697e-6 + -2.7774e-6 * t);
    const double ee = dpsi * std::cos(mean_obliquity_2006(t)) + equinox_complement(t);
    return anp(gmst_2006(ut1, ut2, tt1, tt2) + ee);
}

// ------------------------------------------------------------ geom.rs
// geom.rs:265-268: the MJD is kept as the second part of the date.
void source_vector(double ra, double dec, double mjd, double s[3]) {   // geom.rs:289-297
    const double gst = gast_2006(kMjdZero, mjd, kMjdZero, mjd);
    const double ha = gst - ra;
    const double sd = std::sin(dec), cd = std::cos(dec);
    const double sh = std::sin(ha), ch = std::cos(ha);
    s[0] = cd * ch;
    s[1] = -cd * sh;
    s[2] = sd;
}

double antenna_delay(const double *x, double ra, double dec, double mjd) {   // geom.rs:300-309
    double s[3];
    source_vector(ra, dec, mjd, s);
    return -1.0 * (x[0] * s[0] + x[1] * s[1] + x[2] * s[2]) / kC;
}

double baseline_delay(const double *x1, const double *x2, double ra, double dec, double mjd) {   // geom.rs:313-327
    double s[3];
    source_vector(ra, dec, mjd, s);
    const double bx = x2[0] - x1[0], by = x2[1] - x1[1], bz = x2[2] - x1[2];
    return -1.0 * (bx * s[0] + by * s[1] + bz * s[2]) / kC;
}

// ------------------------------------------------------------ parsing
// Rust's str::parse::<f64>: the whole string, decimal or inf/infinity/nan.
bool parse_f64(const std::string &s, double *out) {
    if (s.empty()) return false;
    size_t i = 0;
    if (s[i] == '+' || s[i] == '-') ++i;
    if (i == s.size()) return false;
    const std::string body = s.substr(i);
    std::string low;
    for (char ch : body) low += (char)std::tolower((unsigned char)ch);
    if (low == "inf" || low == "infinity" || low == "nan") {
        const double v = low == "nan" ? NAN : INFINITY;
        *out = (s[0] == '-') ? -v : v;
        return true;
    }
    bool digits = false, dot = false;
    size_t j = 0;
    for (; j < body.size() && (std::isdigit((unsigned char)body[j]) || body[j] == '.'); ++j) {
        if (body[j] == '.') {
            if (dot) return false;
            dot = true;
        } else {
            digits = true;
        }
    }
    if (!digits) return false;
    if (j < body.size()) {
        if (body[j] != 'e' && body[j] != 'E') return false;
        ++j;
        if (j < body.size() && (body[j] == '+' || body[j] == '-')) ++j;
        if (j == body.size()) return false;
        for (; j < body.size(); ++j)
            if (!std::isdigit((unsigned char)body[j])) return false;
    }
    *out = std::strtod(s.c_str(), nullptr);
    return true;
}

// Rust's str::parse for unsigned (u32) and signed (i32) integers.
bool parse_int(const std::string &s, bool is_signed, int64_t lo, int64_t hi, int64_t *out) {
    size_t i = 0;
    bool neg = false;
    if (!s.empty() && (s[0] == '+' || (is_signed && s[0] == '-'))) {
        neg = s[0] == '-';
        i = 1;
    }
    if (i == s.size()) return false;
    int64_t v = 0;
    for (; i < s.size(); ++i) {
        if (!std::isdigit((unsigned char)s[i])) return false;
        v = v * 10 + (s[i] - '0');
        if (v > hi + 1) return false;
    }
    if (neg) v = -v;
    if (v < lo || v > hi) return false;
    *out = v;
    return true;
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::vector<std::string> split_whitespace(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) ++j;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

bool in_minutes_range(double v) { return v >= 0.0 && v < 60.0; }

// geom.rs:21-31
bool packed_sexagesimal(double value, double hms[3]) {
    const double a = std::fabs(value);
    const double h = std::floor(a / 10000.0);
    const double rem = a - h * 10000.0;
    const double m = std::floor(rem / 100.0);
    const double s = rem - m * 100.0;
    if (!in_minutes_range(m) || !in_minutes_range(s)) return false;
    hms[0] = h;
    hms[1] = m;
    hms[2] = s;
    return true;
}

// geom.rs:33-81: markers become spaces; one field is packed, three are h m s.
bool sexagesimal_fields(std::string text, const char *markers, double hms[3]) {
    for (char &ch : text)
        if (std::strchr(markers, ch)) ch = ' ';
    const std::vector<std::string> parts = split_whitespace(text);
    if (parts.size() == 1) {
        double v;
        return parse_f64(parts[0], &v) && packed_sexagesimal(v, hms);
    }
    if (parts.size() != 3) return false;
    for (int i = 0; i < 3; ++i)
        if (!parse_f64(parts[i], &hms[i])) return false;
    return in_minutes_range(hms[1]) && in_minutes_range(hms[2]);
}

std::string lowercase(std::string s) {
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

// geom.rs:152-184
bool iso_to_mjd(std::string text, double *mjd) {
    for (char &ch : text)
        if (ch == 'z') ch = 'Z';
    std::vector<std::string> parts;
    std::string cur;
    for (char ch : text) {
        if (ch == '-' || ch == 'T' || ch == ':' || ch == 'Z') {
            parts.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    parts.push_back(cur);
    if (parts.size() < 6) return false;
    int64_t year, month, day, hour, minute;
    double second;
    if (!parse_int(parts[0], true, INT32_MIN, INT32_MAX, &year) ||
        !parse_int(parts[1], false, 0, UINT32_MAX, &month) || !parse_int(parts[2], false, 0, UINT32_MAX, &day) ||
        !parse_int(parts[3], false, 0, UINT32_MAX, &hour) || !parse_int(parts[4], false, 0, UINT32_MAX, &minute) ||
        !parse_f64(parts[5], &second))
        return false;
    // Julian day number (Meeus): January and February count as months 13, 14
    // of the previous year; integer division truncates.
    const int64_t y = month <= 2 ? year - 1 : year;
    const int64_t m = month <= 2 ? month + 12 : month;
    const int64_t a = y / 100;
    const int64_t b = 2 - a + a / 4;
    const int64_t jd_int = (int64_t)(int32_t)std::floor(365.25 * (double)(y + 4716)) +
                           (int64_t)(int32_t)std::floor(30.6001 * (double)(m + 1)) + day + b - 1524;
    const double frac = (double)hour / 24.0 + (double)minute / 1440.0 + second / 86400.0;
    const double jd = (double)(int32_t)jd_int - 0.5 + frac;
    *mjd = jd - kMjdZero;
    return true;
}

// geom.rs:186-224
bool year_doy_to_mjd(const std::string &text, double *mjd) {
    const std::vector<std::string> f = split_whitespace(text);
    if (f.empty()) return false;
    const std::string time = f.size() > 1 ? f[1] : "00:00:00";
    const size_t slash = f[0].find('/');
    if (slash == std::string::npos) return false;
    int64_t year, doy;
    if (!parse_int(f[0].substr(0, slash), true, INT32_MIN, INT32_MAX, &year) ||
        !parse_int(f[0].substr(slash + 1), false, 0, UINT32_MAX, &doy))
        return false;
    const bool leap = (year % 4 == 0 && year % 100 != 0) || (year % 400 == 0);
    const int mdays[12] = {31, leap ? 29 : 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    const int64_t max_doy = leap ? 366 : 365;
    if (doy == 0 || doy > max_doy) return false;
    int month = 0;
    int64_t rem = doy;
    while (rem > mdays[month]) rem -= mdays[month++];
    char head[64];
    std::snprintf(head, sizeof head, "%04d-%02d-%02dT", (int)year, month + 1, (int)rem);
    return iso_to_mjd(std::string(head) + time + "Z", mjd);
}

fxc_status bad(const char *what) { return fxc_internal_fail(FXC_ERR_INVALID_ARG, what); }

}  // namespace

extern "C" fxc_status fxc_parse_ra(const char *text, double *radians) {
    if (!text || !radians) return bad("fxc_parse_ra: null argument");
    const std::string raw = lowercase(trim(text));
    if (raw.empty()) return bad("Empty RA");
    const bool marked = raw.find_first_of("hms: ") != std::string::npos;
    double hours;
    if (marked) {
        double f[3];
        if (!sexagesimal_fields(raw, "hms:", f)) return bad("Invalid hms value");
        hours = f[0] + f[1] / 60.0 + f[2] / 3600.0;
    } else {
        double v;
        if (!parse_f64(raw, &v)) return bad("Invalid RA number");
        if (std::fabs(v) >= 10000.0) {
            double f[3];
            if (!packed_sexagesimal(v, f)) return bad("Invalid hhmmss/ddmmss value");
            hours = f[0] + f[1] / 60.0 + f[2] / 3600.0;
        } else {
            *radians = to_radians(v);   // decimal degrees
            return FXC_OK;
        }
    }
    *radians = to_radians(hours * 15.0);
    return FXC_OK;
}

extern "C" fxc_status fxc_parse_dec(const char *text, double *radians) {
    if (!text || !radians) return bad("fxc_parse_dec: null argument");
    const std::string raw = lowercase(trim(text));
    if (raw.empty()) return bad("Empty Dec");
    const double sign = raw[0] == '-' ? -1.0 : 1.0;
    const std::string body = raw.substr(raw.find_first_not_of("+-") == std::string::npos
                                            ? raw.size()
                                            : raw.find_first_not_of("+-"));
    const bool marked = body.find_first_of("dms'\": ") != std::string::npos;
    double degrees;
    if (marked) {
        double f[3];
        if (!sexagesimal_fields(body, "dms'\":", f)) return bad("Invalid dms value");
        degrees = f[0] + f[1] / 60.0 + f[2] / 3600.0;
    } else {
        double v;
        if (!parse_f64(body, &v)) return bad("Invalid Dec number");
        if (std::fabs(v) >= 10000.0) {
            double f[3];
            if (!packed_sexagesimal(v, f)) return bad("Invalid hhmmss/ddmmss value");
            degrees = f[0] + f[1] / 60.0 + f[2] / 3600.0;
        } else {
            *radians = to_radians(sign * v);
            return FXC_OK;
        }
    }
    *radians = to_radians(sign * degrees);
    return FXC_OK;
}

extern "C" fxc_status fxc_parse_epoch_mjd(const char *text, double *mjd) {
    if (!text || !mjd) return bad("fxc_parse_epoch_mjd: null argument");
    const std::string raw = trim(text);
    if (raw.empty()) return bad("Empty epoch");
    if (raw.find('/') != std::string::npos)
        return year_doy_to_mjd(raw, mjd) ? FXC_OK : bad("Invalid YYYY/DDD epoch");
    if (raw.find_first_of("-T:Zz") != std::string::npos)
        return iso_to_mjd(raw, mjd) ? FXC_OK : bad("Invalid epoch format");
    double v;
    const std::string num = (raw[0] == 'J' || raw[0] == 'j') ? raw.substr(1) : raw;
    if (!parse_f64(num, &v)) return bad("Invalid epoch number");
    if (v >= 40000.0 && v < 100000.0) {   // an MJD
        *mjd = v;
        return FXC_OK;
    }
    if (v >= 1800.0 && v < 3000.0) {      // a Julian year
        *mjd = 51544.5 + (v - 2000.0) * 365.25;
        return FXC_OK;
    }
    return bad("Unsupported epoch format. Use ISO datetime, MJD, or year (e.g. 2000)");
}

extern "C" fxc_status fxc_precess_j2000_to_date(double ra, double dec, double mjd, double *ra_date,
                                               double *dec_date) {
    if (!ra_date || !dec_date || !std::isfinite(mjd)) return bad("fxc_precess_j2000_to_date: bad argument");
    Mat3 r;
    bias_precession_2006(kMjdZero, mjd, r);
    const double cp = std::cos(dec);
    const double p[3] = {std::cos(ra) * cp, std::sin(ra) * cp, std::sin(dec)};
    double q[3];
    for (int j = 0; j < 3; ++j) {
        double w = 0.0;
        for (int i = 0; i < 3; ++i) w += r[j][i] * p[i];
        q[j] = w;
    }
    const double d2 = q[0] * q[0] + q[1] * q[1];
    const double th = d2 == 0.0 ? 0.0 : std::atan2(q[1], q[0]);
    const double ph = q[2] == 0.0 ? 0.0 : std::atan2(q[2], std::sqrt(d2));
    *ra_date = anp(th);
    *dec_date = ph;
    return FXC_OK;
}

extern "C" fxc_status fxc_sidereal_time(double mjd, double *gmst, double *gast) {
    if (!std::isfinite(mjd)) return bad("fxc_sidereal_time: bad MJD");
    if (gmst) *gmst = gmst_2006(kMjdZero, mjd, kMjdZero, mjd);
    if (gast) *gast = gast_2006(kMjdZero, mjd, kMjdZero, mjd);
    return FXC_OK;
}

extern "C" fxc_status fxc_geometric_delay(const double ant1_ecef[3], const double ant2_ecef[3], double ra,
                                          double dec, double mjd, fxc_geometric_delay_result *out) {
    if (!ant1_ecef || !ant2_ecef || !out || !std::isfinite(mjd)) return bad("fxc_geometric_delay: bad argument");
    const double dt_s = 1.0;
    const double dt_mjd = dt_s / 86400.0;
    const double minus = baseline_delay(ant1_ecef, ant2_ecef, ra, dec, mjd - dt_mjd);
    out->delay1_s = antenna_delay(ant1_ecef, ra, dec, mjd);
    out->delay2_s = antenna_delay(ant2_ecef, ra, dec, mjd);
    const double now = baseline_delay(ant1_ecef, ant2_ecef, ra, dec, mjd);
    const double plus = baseline_delay(ant1_ecef, ant2_ecef, ra, dec, mjd + dt_mjd);
    out->delay_s = now;
    out->rate_sps = (plus - minus) / (2.0 * dt_s);
    out->accel_sps2 = (plus - 2.0 * now + minus) / (dt_s * dt_s);
    return FXC_OK;
}

extern "C" fxc_status fxc_geometric_delay_table(const double ant1_ecef[3], const double ant2_ecef[3], double ra,
                                                double dec, double mjd, double time_offset_s, int64_t n_points,
                                                double *table) {
    if (!ant1_ecef || !ant2_ecef || !table || n_points < 1 || !std::isfinite(mjd) || !std::isfinite(time_offset_s))
        return bad("fxc_geometric_delay_table: bad argument");
    for (int64_t sec = 0; sec < n_points; ++sec) {
        const double t_abs = time_offset_s + (double)sec;
        table[sec] = baseline_delay(ant1_ecef, ant2_ecef, ra, dec, mjd + t_abs / 86400.0);
    }
    return FXC_OK;
}

extern "C" fxc_status fxc_geometric_model_bound(const double ant1_ecef[3], const double ant2_ecef[3],
                                                double obs_hz, double *delay_s, double *phase_rad) {
    if (!ant1_ecef || !ant2_ecef || !std::isfinite(obs_hz) || obs_hz < 0.0)
        return bad("fxc_geometric_model_bound: bad argument");
    // IAU 2000B vs 2000A nutation: at most 1 mas (published accuracy of the
    // 2000B series, 1995-2050) in the equation of the equinoxes
    const double dtheta = 1e-3 / 3600.0 * (M_PI / 180.0);
    const double bx = ant2_ecef[0] - ant1_ecef[0], by = ant2_ecef[1] - ant1_ecef[1],
                 bz = ant2_ecef[2] - ant1_ecef[2];
    const double d = std::sqrt(bx * bx + by * by + bz * bz) * dtheta / 299792458.0;
    if (delay_s) *delay_s = d;
    if (phase_rad) *phase_rad = 2.0 * M_PI * obs_hz * d;
    return FXC_OK;
}

extern "C" int64_t fxc_geometric_table_points(int64_t total_frames, double frame_dt) {
    const double dur = (double)total_frames * frame_dt;   // main.rs:1630-1634
    const double secs = std::fmax(std::ceil(dur), 1.0);
    return (int64_t)secs + 1;
}
