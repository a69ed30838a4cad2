"""Oracle (test infrastructure only): station geometry of geom.rs and the
1-second geometric-delay table of main.rs:1628-1642, restated in plain
Python for the tests to check the library against.  Nothing in the library
imports this module.

The reference calls ERFA (erfa 0.2.1, absent here).  The ERFA routines on
its path are restated from the published models:

  era      Earth rotation angle, IAU 2000            (ERFA eraEra00)
  gmst     GMST, IAU 2006                             (eraGmst06)
  obl      mean obliquity, IAU 2006                   (eraObl06)
  fw       Fukushima-Williams precession angles 2006  (eraPfw06)
  pmat     bias-precession matrix                     (eraPmat06 = eraFw2m(eraPfw06))
  nut2000b nutation, IAU 2000B, 77 luni-solar terms   (eraNut00b)
  eect     equation-of-equinoxes complement,
           terms above 0.5 microarcsecond             (eraEect00)
  gast     GMST + dpsi cos(eps_A) + eect, dpsi = 2000B with the IAU 2006
           adjustments, standing in for eraGst06a (IAU 2000A nutation)

Pins (tests/test_geom_cpu.py): ERFA's published test values for each
routine (liberfa/erfa test suite); the reference's own tests of parse_ra,
parse_dec, parse_epoch_to_mjd and the GMST rate (geom.rs:387-437); closed
forms (zero baseline, a source at the pole, the analytic rate and
acceleration of a rotating baseline).  The 2000B series as restated here
differs from ERFA's eraNut00b value by 0.115 mas in dpsi at 2006-01-01,
and the apparent sidereal time from eraGst06a by 2.0e-9 rad."""
import math

AS2R = 4.848136811095359935899141e-6
TURNAS = 1296000.0
D2PI = 2.0 * math.pi
DJ00 = 2451545.0
DJC = 36525.0
MJD0 = 2400000.5
C_LIGHT = 299792458.0                      # geom.rs:4


def anp(a):
    w = math.fmod(a, D2PI)
    return w + D2PI if w < 0 else w


def _poly(c, t):
    acc = 0.0
    for k in reversed(c):
        acc = acc * t + k
    return acc


def era(d1, d2):
    """Earth rotation angle: 2 pi (0.7790572732640 + 1.00273781191135448 Tu)."""
    lo, hi = (d1, d2) if d1 < d2 else (d2, d1)
    tu = lo + (hi - DJ00)
    frac = math.fmod(lo, 1.0) + math.fmod(hi, 1.0)
    return anp(D2PI * (frac + 0.7790572732640 + 0.00273781191135448 * tu))


def centuries(d1, d2):
    return ((d1 - DJ00) + d2) / DJC


def gmst(u1, u2, t1, t2):
    t = centuries(t1, t2)
    p = _poly([0.014506, 4612.156534, 1.3915817, -0.00000044, -0.000029956, -0.0000000368], t)
    return anp(era(u1, u2) + p * AS2R)


def obl(t):
    return _poly([84381.406, -46.836769, -0.0001831, 0.00200340, -0.000000576, -0.0000000434], t) * AS2R


def fw(d1, d2):
    """(gamma_bar, phi_bar, psi_bar, eps_A) in radians."""
    t = centuries(d1, d2)
    g = _poly([-0.052928, 10.556378, 0.4932044, -0.00031238, -0.000002788, 0.0000000260], t) * AS2R
    ph = _poly([84381.412819, -46.811016, 0.0511268, 0.00053289, -0.000000440, -0.0000000176], t) * AS2R
    ps = _poly([-0.041775, 5038.481484, 1.5584175, -0.00018522, -0.000026452, -0.0000000148], t) * AS2R
    return g, ph, ps, obl(t)


def _rz(a):
    s, c = math.sin(a), math.cos(a)
    return [[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]]


def _rx(a):
    s, c = math.sin(a), math.cos(a)
    return [[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]]


def _mm(a, b):
    return [[a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] for j in range(3)]
            for i in range(3)]


def pmat(d1, d2):
    """R = Rx(-eps_A) Rz(-psi) Rx(phi) Rz(gamma)."""
    g, ph, ps, ep = fw(d1, d2)
    return _mm(_rx(-ep), _mm(_rz(-ps), _mm(_rx(ph), _rz(g))))


# IAU 2000B luni-solar terms: (l, l', F, D, Om), dpsi sin, sin*t, cos;
# deps cos, cos*t, sin; units 0.1 microarcsecond.
NUT2000B = [
    ((0, 0, 0, 0, 1), -172064161, -174666, 33386, 92052331, 9086, 15377),
    ((0, 0, 2, -2, 2), -13170906, -1675, -13696, 5730336, -3015, -4587),
    ((0, 0, 2, 0, 2), -2276413, -234, 2796, 978459, -485, 1374),
    ((0, 0, 0, 0, 2), 2074554, 207, -698, -897492, 470, -291),
    ((0, 1, 0, 0, 0), 1475877, -3633, 11817, 73871, -184, -1924),
    ((0, 1, 2, -2, 2), -516821, 1226, -524, 224386, -677, -174),
    ((1, 0, 0, 0, 0), 711159, 73, -872, -6750, 0, 358),
    ((0, 0, 2, 0, 1), -387298, -367, 380, 200728, 18, 318),
    ((1, 0, 2, 0, 2), -301461, -36, 816, 129025, -63, 367),
    ((0, -1, 2, -2, 2), 215829, -494, 111, -95929, 299, 132),
    ((0, 0, 2, -2, 1), 128227, 137, 181, -68982, -9, 39),
    ((-1, 0, 2, 0, 2), 123457, 11, 19, -53311, 32, -4),
    ((-1, 0, 0, 2, 0), 156994, 10, -168, -1235, 0, 82),
    ((1, 0, 0, 0, 1), 63110, 63, 27, -33228, 0, -9),
    ((-1, 0, 0, 0, 1), -57976, -63, -189, 31429, 0, -75),
    ((-1, 0, 2, 2, 2), -59641, -11, 149, 25543, -11, 66),
    ((1, 0, 2, 0, 1), -51613, -42, 129, 26366, 0, 78),
    ((-2, 0, 2, 0, 1), 45893, 50, 31, -24236, -10, 20),
    ((0, 0, 0, 2, 0), 63384, 11, -150, -1220, 0, 29),
    ((0, 0, 2, 2, 2), -38571, -1, 158, 16452, -11, 68),
    ((0, -2, 2, -2, 2), 32481, 0, 0, -13870, 0, 0),
    ((-2, 0, 0, 2, 0), -47722, 0, -18, 477, 0, -25),
    ((2, 0, 2, 0, 2), -31046, -1, 131, 13238, -11, 59),
    ((1, 0, 2, -2, 2), 28593, 0, -1, -12338, 10, -3),
    ((-1, 0, 2, 0, 1), 20441, 21, 10, -10758, 0, -3),
    ((2, 0, 0, 0, 0), 29243, 0, -74, -609, 0, 13),
    ((0, 0, 2, 0, 0), 25887, 0, -66, -550, 0, 11),
    ((0, 1, 0, 0, 1), -14053, -25, 79, 8551, -2, -45),
    ((-1, 0, 0, 2, 1), 15164, 10, 11, -8001, 0, -1),
    ((0, 2, 2, -2, 2), -15794, 72, -16, 6850, -42, -5),
    ((0, 0, -2, 2, 0), 21783, 0, 13, -167, 0, 13),
    ((1, 0, 0, -2, 1), -12873, -10, -37, 6953, 0, -14),
    ((0, -1, 0, 0, 1), -12654, 11, 63, 6415, 0, 26),
    ((-1, 0, 2, 2, 1), -10204, 0, 25, 5222, 0, 15),
    ((0, 2, 0, 0, 0), 16707, -85, -10, 168, -1, 10),
    ((1, 0, 2, 2, 2), -7691, 0, 44, 3268, 0, 19),
    ((-2, 0, 2, 0, 0), -11024, 0, -14, 104, 0, 2),
    ((0, 1, 2, 0, 2), 7566, -21, -11, -3250, 0, -5),
    ((0, 0, 2, 2, 1), -6637, -11, 25, 3353, 0, 14),
    ((0, -1, 2, 0, 2), -7141, 21, 8, 3070, 0, 4),
    ((0, 0, 0, 2, 1), -6302, -11, 2, 3272, 0, 4),
    ((1, 0, 2, -2, 1), 5800, 10, 2, -3045, 0, -1),
    ((2, 0, 2, -2, 2), 6443, 0, -7, -2768, 0, -4),
    ((-2, 0, 0, 2, 1), -5774, -11, -15, 3041, 0, -5),
    ((2, 0, 2, 0, 1), -5350, 0, 21, 2695, 0, 12),
    ((0, -1, 2, -2, 1), -4752, -11, -3, 2719, 0, -3),
    ((0, 0, 0, -2, 1), -4940, -11, -21, 2720, 0, -9),
    ((-1, -1, 0, 2, 0), 7350, 0, -8, -51, 0, 4),
    ((2, 0, 0, -2, 1), 4065, 0, 6, -2206, 0, 1),
    ((1, 0, 0, 2, 0), 6579, 0, -24, -199, 0, 2),
    ((0, 1, 2, -2, 1), 3579, 0, 5, -1900, 0, 1),
    ((1, -1, 0, 0, 0), 4725, 0, -6, -41, 0, 3),
    ((-2, 0, 2, 0, 2), -3075, 0, -2, 1313, 0, -1),
    ((3, 0, 2, 0, 2), -2904, 0, 15, 1233, 0, 7),
    ((0, -1, 0, 2, 0), 4348, 0, -10, -81, 0, 2),
    ((1, -1, 2, 0, 2), -2878, 0, 8, 1232, 0, 4),
    ((0, 0, 0, 1, 0), -4230, 0, 5, -20, 0, -2),
    ((-1, -1, 2, 2, 2), -2819, 0, 7, 1207, 0, 3),
    ((-1, 0, 2, 0, 0), -4056, 0, 5, 40, 0, -2),
    ((0, -1, 2, 2, 2), -2647, 0, 11, 1129, 0, 5),
    ((-2, 0, 0, 0, 1), -2294, 0, -10, 1266, 0, -4),
    ((1, 1, 2, 0, 2), 2481, 0, -7, -1062, 0, -3),
    ((2, 0, 0, 0, 1), 2179, 0, -2, -1129, 0, -2),
    ((-1, 1, 0, 1, 0), 3276, 0, 1, -9, 0, 0),
    ((1, 1, 0, 0, 0), -3389, 0, 5, 35, 0, -2),
    ((1, 0, 2, 0, 0), 3339, 0, -13, -107, 0, 1),
    ((-1, 0, 2, -2, 1), -1987, 0, -6, 1073, 0, -2),
    ((1, 0, 0, 0, 2), -1981, 0, 0, 854, 0, 0),
    ((-1, 0, 0, 1, 0), 4026, 0, -353, -553, 0, -139),
    ((0, 0, 2, 1, 2), 1660, 0, -5, -710, 0, -2),
    ((-1, 0, 2, 4, 2), -1521, 0, 9, 647, 0, 4),
    ((-1, 1, 0, 1, 1), 1314, 0, 0, -700, 0, 0),
    ((0, -2, 2, -2, 1), -1283, 0, 0, 672, 0, 0),
    ((1, 2, 2, -2, 2), -1331, 0, 8, 663, 0, 4),
    ((-2, 0, 2, 2, 2), 1383, 0, -2, -594, 0, -2),
    ((-1, 0, 0, 0, 2), 1405, 0, 4, -610, 0, 2),
    ((1, 1, 2, -2, 2), 1290, 0, 0, -556, 0, 0),
]


def nut2000b(d1, d2):
    """(dpsi, deps) in radians, IAU 2000B: linear Delaunay arguments, the
    77 terms summed smallest first, plus the fixed planetary offsets
    -0.135 mas and +0.388 mas."""
    t = centuries(d1, d2)
    lin = [(485868.249036, 1717915923.2178), (1287104.79305, 129596581.0481),
           (335779.526232, 1739527262.8478), (1072260.70369, 1602961601.2090),
           (450160.398036, -6962890.5431)]
    fa = [math.fmod(a + b * t, TURNAS) * AS2R for a, b in lin]
    dp = de = 0.0
    for n, ps, pst, pc, ec, ect, es in reversed(NUT2000B):
        arg = math.fmod(sum(k * x for k, x in zip(n, fa)), D2PI)
        sa, ca = math.sin(arg), math.cos(arg)
        dp += (ps + pst * t) * sa + pc * ca
        de += (ec + ect * t) * ca + es * sa
    u = AS2R / 1e7
    return dp * u + -0.135 * (AS2R / 1e3), de * u + 0.388 * (AS2R / 1e3)


def delaunay(t):
    """Delaunay arguments l, l', F, D, Om (IERS Conventions 2003), radians."""
    cs = [[485868.249036, 1717915923.2178, 31.8792, 0.051635, -0.00024470],
          [1287104.793048, 129596581.0481, -0.5532, 0.000136, -0.00001149],
          [335779.526232, 1739527262.8478, -12.7512, -0.001037, 0.00000417],
          [1072260.703692, 1602961601.2090, -6.3706, 0.006593, -0.00003169],
          [450160.398036, -6962890.5431, 7.4722, 0.007702, -0.00005939]]
    return [math.fmod(_poly(c, t), TURNAS) * AS2R for c in cs]


EECT0 = [((0, 0, 0, 0, 1), 2640.96e-6, -0.39e-6), ((0, 0, 0, 0, 2), 63.52e-6, -0.02e-6),
         ((0, 0, 2, -2, 3), 11.75e-6, 0.01e-6), ((0, 0, 2, -2, 1), 11.21e-6, 0.01e-6),
         ((0, 0, 2, -2, 2), -4.55e-6, 0.0), ((0, 0, 2, 0, 3), 2.02e-6, 0.0),
         ((0, 0, 2, 0, 1), 1.98e-6, 0.0), ((0, 0, 0, 0, 3), -1.72e-6, 0.0),
         ((0, 1, 0, 0, 1), -1.41e-6, -0.01e-6), ((0, 1, 0, 0, -1), -1.26e-6, -0.01e-6),
         ((1, 0, 0, 0, -1), -0.63e-6, 0.0), ((1, 0, 0, 0, 1), -0.63e-6, 0.0)]


def eect(d1, d2):
    t = centuries(d1, d2)
    fa = delaunay(t)
    s0 = 0.0
    for n, s, c in reversed(EECT0):
        arg = sum(k * x for k, x in zip(n, fa))
        s0 += s * math.sin(arg) + c * math.cos(arg)
    s1 = -0.87e-6 * math.sin(fa[4])
    return (s0 + s1 * t) * AS2R


def gast(u1, u2, t1, t2):
    t = centuries(t1, t2)
    dp, _ = nut2000b(t1, t2)
    dpsi = dp + dp * (0.4697e-6 + -2.7774e-6 * t)      # IAU 2006 adjustments
    ee = dpsi * math.cos(obl(t)) + eect(t1, t2)         # equation of the equinoxes
    return anp(gmst(u1, u2, t1, t2) + ee)


# ------------------------------------------------------------------ geom.rs
def precess(ra, dec, mjd):
    """precess_j2000_to_mean_of_date (geom.rs:270-277)."""
    r = pmat(MJD0, mjd)
    cp = math.cos(dec)
    p = [math.cos(ra) * cp, math.sin(ra) * cp, math.sin(dec)]
    q = [sum(r[j][i] * p[i] for i in range(3)) for j in range(3)]
    d2 = q[0] * q[0] + q[1] * q[1]
    th = 0.0 if d2 == 0.0 else math.atan2(q[1], q[0])
    ph = 0.0 if q[2] == 0.0 else math.atan2(q[2], math.sqrt(d2))
    return anp(th), ph


def sidereal(mjd):
    """(GMST, GAST) for an MJD used as both UT1 and TT (geom.rs:282-297)."""
    return gmst(MJD0, mjd, MJD0, mjd), gast(MJD0, mjd, MJD0, mjd)


def source_vector(ra, dec, mjd):                 # geom.rs:289-297
    ha = gast(MJD0, mjd, MJD0, mjd) - ra
    return [math.cos(dec) * math.cos(ha), -math.cos(dec) * math.sin(ha), math.sin(dec)]


def baseline_delay(x1, x2, ra, dec, mjd):        # geom.rs:313-327
    s = source_vector(ra, dec, mjd)
    b = [x2[i] - x1[i] for i in range(3)]
    return -1.0 * (b[0] * s[0] + b[1] * s[1] + b[2] * s[2]) / C_LIGHT


def antenna_delay(x, ra, dec, mjd):              # geom.rs:300-309
    s = source_vector(ra, dec, mjd)
    return -1.0 * (x[0] * s[0] + x[1] * s[1] + x[2] * s[2]) / C_LIGHT


def geometric_delay(x1, x2, ra, dec, mjd):
    """calculate_geometric_delay_and_derivatives (geom.rs:346-384):
    (delay1, delay2, delay, rate, accel)."""
    h = 1.0 / 86400.0
    m, z, p = (baseline_delay(x1, x2, ra, dec, mjd + k * h) for k in (-1, 0, 1))
    return (antenna_delay(x1, ra, dec, mjd), antenna_delay(x2, ra, dec, mjd), z,
            (p - m) / 2.0, (p - 2.0 * z + m) / 1.0)


def delay_table(x1, x2, ra, dec, mjd, time_offset_s, n_points):
    """main.rs:1628-1642."""
    return [baseline_delay(x1, x2, ra, dec, mjd + (time_offset_s + k) / 86400.0)
            for k in range(n_points)]


def table_points(total_frames, frame_dt):
    return int(max(math.ceil(total_frames * frame_dt), 1.0)) + 1


# ------------------------------------------------------------- parsing
def _f64(s):
    """Rust's str::parse::<f64> on the strings these parsers produce."""
    import re
    if not re.fullmatch(r"[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|inf|infinity|nan)", s, re.I):
        raise ValueError(s)
    return float(s)


def _int(s, signed):
    import re
    if not re.fullmatch(r"[+-]?\d+" if signed else r"\+?\d+", s):
        raise ValueError(s)
    return int(s)


def _floor(x):                                   # f64::floor: inf and nan pass through
    return float(math.floor(x)) if math.isfinite(x) else x


def _packed(v):                                  # geom.rs:21-31
    a = abs(v)
    h = _floor(a / 10000.0)
    rem = a - h * 10000.0
    m = _floor(rem / 100.0)
    s = rem - m * 100.0
    if not (0.0 <= m < 60.0 and 0.0 <= s < 60.0):
        raise ValueError("packed")
    return h, m, s


def _fields(text, markers):                      # geom.rs:33-81
    for ch in markers:
        text = text.replace(ch, " ")
    parts = text.split()
    if len(parts) == 1:
        return _packed(_f64(parts[0]))
    if len(parts) != 3:
        raise ValueError("fields")
    h, m, s = (_f64(x) for x in parts)
    if not (0.0 <= m < 60.0 and 0.0 <= s < 60.0):
        raise ValueError("range")
    return h, m, s


def _rad(deg):
    return deg * (math.pi / 180.0)


def parse_ra(text):                              # geom.rs:86-111
    raw = text.strip().lower()
    if not raw:
        raise ValueError("empty")
    if any(c in raw for c in "hms: "):
        h, m, s = _fields(raw, "hms:")
    else:
        v = _f64(raw)
        if not abs(v) >= 10000.0:                # nan takes this branch too
            return _rad(v)
        h, m, s = _packed(v)
    return _rad((h + m / 60.0 + s / 3600.0) * 15.0)


def parse_dec(text):                             # geom.rs:115-148
    raw = text.strip().lower()
    if not raw:
        raise ValueError("empty")
    sign = -1.0 if raw.startswith("-") else 1.0
    body = raw.lstrip("+-")
    if any(c in body for c in "dms'\": "):
        d, m, s = _fields(body, "dms'\":")
    else:
        v = _f64(body)
        if not abs(v) >= 10000.0:
            return _rad(sign * v)
        d, m, s = _packed(v)
    return _rad(sign * (d + m / 60.0 + s / 3600.0))


def _iso(text):                                  # geom.rs:152-184
    import re
    parts = re.split(r"[-T:Z]", text.replace("z", "Z"))
    if len(parts) < 6:
        raise ValueError("iso")
    year = _int(parts[0], True)
    month, day, hour, minute = (_int(p, False) for p in parts[1:5])
    second = _f64(parts[5])
    y, m = (year - 1, month + 12) if month <= 2 else (year, month)
    a = int(y / 100)                             # truncating division
    b = 2 - a + int(a / 4)
    jd_int = math.floor(365.25 * (y + 4716)) + math.floor(30.6001 * (m + 1)) + day + b - 1524
    frac = hour / 24.0 + minute / 1440.0 + second / 86400.0
    return float(jd_int) - 0.5 + frac - MJD0


def parse_epoch_mjd(text):                       # geom.rs:186-257
    raw = text.strip()
    if not raw:
        raise ValueError("empty")
    if "/" in raw:
        f = raw.split()
        time = f[1] if len(f) > 1 else "00:00:00"
        ys, ds = f[0].split("/", 1)
        year, doy = _int(ys, True), _int(ds, False)
        leap = (year % 4 == 0 and year % 100 != 0) or year % 400 == 0
        md = [31, 29 if leap else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
        if doy == 0 or doy > sum(md):
            raise ValueError("doy")
        mon = 0
        while doy > md[mon]:
            doy -= md[mon]
            mon += 1
        return _iso("%04d-%02d-%02dT%sZ" % (year, mon + 1, doy, time))
    if any(c in raw for c in "-T:Zz"):
        return _iso(raw)
    v = _f64(raw[1:] if raw[:1] in ("J", "j") else raw)
    if 40000.0 <= v < 100000.0:
        return v
    if 1800.0 <= v < 3000.0:
        return 51544.5 + (v - 2000.0) * 365.25
    raise ValueError("epoch")
