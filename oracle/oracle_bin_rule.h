/*
 * Oracle (test infrastructure only): the phase_bin rule of
 * hex_phase/include/ice_crystal/phase_math.hpp:37-81 after the unit checks,
 * written once over BREAL so that the operation census (census.cpp) can count
 * its arithmetic; oracle_hex.c instantiates it with BREAL = double, where every
 * cast below is the identity and the operations are the rule's, in its order.
 *
 * The including file defines BREAL, BFN(name), BVAL(x) (the double value of a
 * BREAL), and the libm entry points BHYPOT, BATAN2, BFLOOR, BSQRT, BFABS,
 * BFMIN, BFMAX; orc_bin, PI_D and DBL_EPSILON come from the includer.
 */

/* (x, y, z): the direction's coordinates in the (e0, e1, ki) frame; n their norm. */
static orc_bin BFN(bin_core_)(const double* ue, uint32_t nt, uint32_t np, BREAL x, BREAL y, BREAL z, BREAL n)
{
    orc_bin b;
    memset(&b, 0, sizeof b);
    const BREAL mu = z / n, s = BHYPOT(x, y) / n;
    b.mu = BVAL(mu);
    if (s <= (BREAL)(256 * DBL_EPSILON)) {
        b.pole = mu >= (BREAL)0 ? 1 : -1;
        b.valid = 1;
        return b;
    }
    const BREAL u = mu >= (BREAL)0 ? ((BREAL)0.5 * s * s) / ((BREAL)1 + mu)
                                   : (BREAL)1 - ((BREAL)0.5 * s * s) / ((BREAL)1 - mu);
    if (!(u >= (BREAL)0 && u <= (BREAL)1)) return b;
    uint32_t lo = 0, hi = nt;
    while (lo + 1 < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (u >= (BREAL)ue[mid]) lo = mid; else hi = mid;
    }
    for (unsigned cand = 0; cand < 2; ++cand) {
        const uint32_t e = lo + cand;
        if (e == 0 || e >= nt) continue;
        const BREAL edge = (BREAL)ue[e];
        const BREAL width = BFMIN(edge - (BREAL)ue[e - 1], (BREAL)ue[e + 1] - edge);
        const BREAL tol = BFMIN((BREAL)(64 * DBL_EPSILON) * BFMAX(BFABS(u), BFABS(edge)), (BREAL)1e-7 * width);
        if (BFABS(u - edge) <= tol) {
            b.snapped = (u != edge);
            lo = e;
            break;
        }
    }
    BREAL phi = BATAN2(y, x);
    if (phi >= (BREAL)PI_D) phi = phi - (BREAL)(2 * PI_D);
    const BREAL v = (phi + (BREAL)PI_D) / (BREAL)(2 * PI_D);
    BREAL index = v * (BREAL)np;
    const BREAL nearest = BFLOOR(index + (BREAL)0.5);
    if (BFABS(index - nearest) <= BFMIN((BREAL)(64 * DBL_EPSILON) * (BREAL)np, (BREAL)1e-7)) {
        b.snapped = b.snapped || (index != nearest);
        index = nearest;
    }
    uint32_t j = (uint32_t)BVAL(index);
    if (j == np) j = 0;
    if (j >= np) return b;
    b.index = (uint64_t)lo * np + j;
    b.valid = 1;
    return b;
}

/* Binning of a direction that need not be unit length (the FP32 drop samples
 * promoted to double): frame projections, their norm, then the rule. */
static orc_bin BFN(bin_raw_)(const double* ue, uint32_t nt, uint32_t np, const orc_frame* g, const double dir[3])
{
    orc_bin b;
    memset(&b, 0, sizeof b);
    const BREAL dx = (BREAL)dir[0], dy = (BREAL)dir[1], dz = (BREAL)dir[2];
    if (!ue || !nt || !np || !(BFABS(dx) <= (BREAL)DBL_MAX && BFABS(dy) <= (BREAL)DBL_MAX && BFABS(dz) <= (BREAL)DBL_MAX))
        return b;
    const BREAL x = (BREAL)g->e0[0] * dx + (BREAL)g->e0[1] * dy + (BREAL)g->e0[2] * dz;
    const BREAL y = (BREAL)g->e1[0] * dx + (BREAL)g->e1[1] * dy + (BREAL)g->e1[2] * dz;
    const BREAL z = (BREAL)g->ki[0] * dx + (BREAL)g->ki[1] * dy + (BREAL)g->ki[2] * dz;
    const BREAL n = BSQRT(x * x + y * y + z * z);
    if (!(n > (BREAL)0) || !(BFABS(n) <= (BREAL)DBL_MAX)) return b;
    return BFN(bin_core_)(ue, nt, np, x, y, z, n);
}
