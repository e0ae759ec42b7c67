"""The matrix entry points of the device boundary against the reference's
own golden vectors (src/matrices/*.t.cpp), through the C ABI:
shx_block_apply for ccs::matrix::{circulant, dense, inner_block, block}
(inner_block.t.cpp's identity and random-boundary cases included)
(one derivative operator whose segments hold the matrices) and
shx_csr_create / shx_csr_apply for ccs::matrix::csr.  The vectors were
printed by the reference; the comparisons allow 1e-12 relative.

The reference's block test also checks the integer layout its builder emits
(block.t.cpp:244-398); that layout has no counterpart here (segments carry
positions and a de-duplicated pool instead), so it is not restated."""
import ctypes as C

import numpy as np
import pytest

import shoccs_amd as S
from shoccs_amd._lib import ShxLineBlock, lib

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)


class DerivDesc(C.Structure):  # include/shoccs_hip.h shx_derivative_desc
    _fields_ = [("dir", C.c_int32), ("n", C.c_int32 * 3), ("nR", C.c_int64 * 3), ("width", C.c_int32),
                ("interior", C.c_double * 9), ("blocks", C.POINTER(ShxLineBlock)), ("nblocks", C.c_int64),
                ("pool", dp), ("npool", C.c_int64), ("neu_row", i64p), ("neu_col", i64p), ("neu_w", dp),
                ("nneu", C.c_int64), ("nrrow", C.c_int64), ("rset", i32p), ("rrow", i64p), ("rptr", i64p),
                ("rcol", i64p), ("rw", dp)]


def _dev(ctx, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert lib().shx_alloc(ctx.p, max(a.nbytes, 8), C.byref(p)) == 0
    assert lib().shx_h2d(ctx.p, p, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
    return p


def _host(ctx, p, n):
    out = np.empty(n)
    assert lib().shx_d2h(ctx.p, out.ctypes.data_as(C.c_void_p), p, out.nbytes) == 0
    return out


def _seg(lo, nl, ni, nr, lcols=0, lcol0=0, lcoef=0, rcols=0, rcol0=0, rcoef=0):
    b = ShxLineBlock()
    for k, v in dict(pair=0, lo=lo, nl=nl, ni=ni, nr=nr, lcols=lcols, lcol0=lcol0, lcoef=lcoef, rcols=rcols,
                     rcol0=rcol0, rcoef=rcoef, lobj=-1, lbw=0, robj=-1, rbw=0, flags=0).items():
        setattr(b, k, v)
    return b


def _block_apply(ctx, length, segs, pool, interior, x, b0, accumulate):
    """One line of `length` points along z holding the segments; b = A x
    (accumulate 0) or b0 + A x (1) through shx_block_apply."""
    blocks = (ShxLineBlock * len(segs))(*segs)
    pool = np.ascontiguousarray(list(pool) + [0.0], dtype=np.float64)
    d = DerivDesc()
    d.dir = 2
    d.n[0], d.n[1], d.n[2] = 1, 1, length
    d.width = len(interior)
    for m, c in enumerate(interior):
        d.interior[m] = c
    d.blocks, d.nblocks = blocks, len(segs)
    d.pool, d.npool = pool.ctypes.data_as(dp), pool.size
    op = C.c_void_p()
    rc = lib().shx_derivative_create(ctx.p, C.byref(d), C.byref(op))
    assert rc == 0, lib().shx_last_error(ctx.p)
    px, pb = _dev(ctx, x), _dev(ctx, b0)
    assert lib().shx_block_apply(op, px, pb, accumulate) == 0
    got = _host(ctx, pb, len(b0))
    lib().shx_free(ctx.p, px)
    lib().shx_free(ctx.p, pb)
    lib().shx_derivative_destroy(op)
    return got


def _close(got, want):
    want = np.asarray(want)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want))), (got, want)


def test_circulant_random(ctx):
    """circulant.t.cpp:70-101: circulant{5 rows, row offset 3, stride 1}."""
    coeffs = [-1.8200787083110566, -4.475169398045676, 0.1576129649934348]
    x = [-3.1777625401858884, -1.1637081452907765, -8.784390495345235, -6.1173200419288705, 6.419770261059497,
         3.22476867262354, 4.579902923667724, -7.075669447054981, 2.9684424313172144, -5.8732966768985975,
         3.6343912595600116]
    want = [0.0, 0.0, 0.0, 44.37616458118171, -17.087290102627783, -25.394021164722776, -27.480391451152123,
            23.79690059566026]
    got = _block_apply(ctx, 11, [_seg(3, 0, 5, 0)], [], coeffs, x, np.zeros(11), 0)
    _close(got[:8], want)
    assert np.all(got[8:] == 0.0)


@pytest.mark.parametrize("case", ["non_square1", "non_square2"])
def test_dense(ctx, case):
    """dense.t.cpp:73-135: 5 x 7 and 9 x 3 matrices (one closure block of the
    segment, rows from 0, columns from 0)."""
    if case == "non_square1":
        A = [83.95087197745517, 54.60811141907641, -31.02615838268912, -23.749193978143865, -47.73731285639923,
             68.77294874944397, -84.32582368924861, -68.68982276024278, -66.25620963886223, -14.145795257191367,
             -20.66243053847711, 26.12372414082023, -84.95259217482109, -26.51110805056817, 63.255420433843824,
             -62.63678283082555, -44.15912219148467, 79.63870038246813, 40.13227588220241, 4.250640715510485,
             27.337743054200416, 64.34395968254239, -49.2229980208256, -28.064364905871287, 45.086404615059905,
             -74.58586464667385, 45.51639082665275, 10.633368817791506, 27.62738451878431, -30.951742698251962,
             88.544956462669, 99.34593385752356, 2.3159193728828313, -75.82935876599558, 10.567250113100783]
        r, c = 5, 7
        x = [91.37999616851539, 50.64693725996551, -64.60583179621301, -21.187021053699596, 61.80093732179881,
             -51.51860192935814, 97.12696476389027]
        want = [-1738.7987608832982, -4864.707653579187, 8690.006071322056, -1676.9985963825454, -1792.272168705312]
    else:
        A = [-62.96028312265989, -34.222115425099105, -2.9862605513470157, 23.06985790293345, -76.73747513448777,
             0.9629408474623347, -40.620449988213124, -55.87954317817193, 21.40997482658412, 37.708749410667,
             -99.44892107309963, 57.76722097490671, -32.31615638733223, -16.107987296727003, 3.2555572175240286,
             92.33047063004568, -73.78219097339218, -99.4087327368714, -47.41071272784896, 67.99839324859943,
             -56.75967760541846, -30.570447576312347, -63.39035198957018, -24.724532373486397, 43.50746274626897,
             24.831628943793476, -67.23007810345939]
        r, c = 9, 3
        x = [-71.9024502289571, 70.422079255705, -3.7031798936135942]
        want = [2128.0647588947263, -7066.357808643435, -1093.7287232110582, -9928.67369062508, 1177.1969541409073,
                -11466.553949160041, 8407.725947723382, -2174.430915359314, -1130.6331596949249]
    n = max(r, c)
    xx = np.zeros(n)
    xx[:c] = x
    got = _block_apply(ctx, n, [_seg(0, r, 0, 0, lcols=c, lcol0=0, lcoef=0)], A, [0.0, 0.0, 0.0], xx,
                       np.zeros(n), 0)
    _close(got[:r], want)


def test_block_random_boundary(ctx):
    """block.t.cpp:84-171: two inner blocks (16 columns each, at offsets 1 and
    21, stride 1) of [dense 4 x 5 | circulant 10 rows | dense 2 x 3] on one
    line of 37 points; points 0 and 17-20 hold no row."""
    left = [2.247323503221594, -5.0275337061235135, -0.9845370624957113, 9.621361506824222, -1.4847274599487292,
            -5.920356924707782, -2.5821362891416157, 3.4175272453709056, 8.530925204174977, 6.082313908057927,
            -6.47449633704835, 3.2352002101687702, 4.758308560669512, 1.598859246390333, 0.9341660471574365,
            -2.405757184769108, -1.0503192473307266, 2.485464032420346, 9.816354187065464, 5.090921023204807]
    interior = [0.7763383518842382, 2.749405412018632, -0.7966643678385346]
    right = [-0.12697902221008128, 1.4021179149102307, 0.30625420155821903, 1.288208276267654,
             1.3345939131355147, -1.4713774812532359]
    rhs = [1.1489955608128035, -9.641815125514444, 0.5975150739415511, 5.321795555035614, -5.467954909319733,
           1.4687325444642383, 8.370366165425736, -2.810816575553474, 1.815504174076466, 1.1769187743854737,
           -8.870585410511126, -0.14181686526567816, 7.93970270970086, 0.14183234352470464, -8.054425087325441,
           -3.809075460321715]
    exact = [109.78978122898833, 32.2780625869145, -32.38838457188155, 33.25158538253831, -12.072197714155793,
             -6.87521436567669, 26.39304084900066, -2.6761855166330424, 1.8718030426410628, 11.712151684565809,
             -23.362167910509047, -13.601765954771285, 21.606370954127627, 12.97052379948305, -12.477808805315766,
             -4.962089239867884]
    x = np.array([0.0] + rhs + [0.0] * 4 + rhs)
    want = np.array([0.0] + exact + [0.0] * 4 + exact)
    # inner_block.cpp:17-40: left rows from the offset, interior after them,
    # right rows last with columns from offset + columns - right columns
    segs = [_seg(off, 4, 10, 2, lcols=5, lcol0=off, lcoef=0, rcols=3, rcol0=off + 16 - 3, rcoef=20)
            for off in (1, 21)]
    got = _block_apply(ctx, 37, segs, left + right, interior, x, np.zeros(37), 0)
    _close(got, want)
    # plus_eq adds to what b holds
    b0 = np.arange(37.0)
    got = _block_apply(ctx, 37, segs, left + right, interior, x, b0, 1)
    _close(got, want + b0)


def test_block_identity(ctx):
    """block.t.cpp:31-80: three identity inner blocks (rows 1-16, 17-32,
    34-49) reproduce x except at the points without a row; plus_eq doubles."""
    left = np.eye(4).ravel().tolist()
    right = np.eye(2).ravel().tolist()
    segs = [_seg(off, 4, 10, 2, lcols=4, lcol0=off, lcoef=0, rcols=2, rcol0=off + 16 - 2, rcoef=16)
            for off in (1, 17, 34)]
    x = np.random.default_rng(3).uniform(-1, 1, 50)
    got = _block_apply(ctx, 50, segs, left + right, [0.0, 1.0, 0.0], x, np.zeros(50), 0)
    xx = x.copy()
    xx[0] = xx[33] = 0.0
    assert np.array_equal(got, xx)
    got = _block_apply(ctx, 50, segs, left + right, [0.0, 1.0, 0.0], x, xx, 1)
    assert np.array_equal(got, 2 * xx)


def test_inner_block_identity(ctx):
    """inner_block.t.cpp:31-97 "Identity": [I4 | circulant {1} | I2] square
    (16 rows and columns; the one-coefficient circulant as the stencil
    {0, 1, 0}) gives b = x, and plus_eq 2x; the non-square block
    (row offset 1: a shifted 4 x 5 identity, 9 circulant rows, I2 on the
    last two columns) gives no row at point 0 and x elsewhere."""
    x = np.random.default_rng(5).uniform(-1, 1, 16)
    seg = _seg(0, 4, 10, 2, lcols=4, lcol0=0, lcoef=0, rcols=2, rcol0=14, rcoef=16)
    pool = np.eye(4).ravel().tolist() + np.eye(2).ravel().tolist()
    got = _block_apply(ctx, 16, [seg], pool, [0.0, 1.0, 0.0], x, np.zeros(16), 0)
    assert np.array_equal(got, x)
    got = _block_apply(ctx, 16, [seg], pool, [0.0, 1.0, 0.0], x, got, 1)
    assert np.array_equal(got, 2 * x)
    shifted = np.zeros((4, 5))
    for r in range(4):
        shifted[r, r + 1] = 1.0
    seg = _seg(1, 4, 9, 2, lcols=5, lcol0=0, lcoef=0, rcols=2, rcol0=14, rcoef=20)
    pool = shifted.ravel().tolist() + np.eye(2).ravel().tolist()
    got = _block_apply(ctx, 16, [seg], pool, [0.0, 1.0, 0.0], x, np.zeros(16), 0)
    assert got[0] == 0.0
    assert np.array_equal(got[1:], x[1:])


def test_inner_block_random_boundary(ctx):
    """inner_block.t.cpp:99-168: [dense 4 x 5 | circulant 10 rows | dense
    2 x 3] with row and column offset 2 on 18 points (the first two without
    a row)."""
    left = [2.247323503221594, -5.0275337061235135, -0.9845370624957113, 9.621361506824222, -1.4847274599487292,
            -5.920356924707782, -2.5821362891416157, 3.4175272453709056, 8.530925204174977, 6.082313908057927,
            -6.47449633704835, 3.2352002101687702, 4.758308560669512, 1.598859246390333, 0.9341660471574365,
            -2.405757184769108, -1.0503192473307266, 2.485464032420346, 9.816354187065464, 5.090921023204807]
    interior = [0.7763383518842382, 2.749405412018632, -0.7966643678385346]
    right = [-0.12697902221008128, 1.4021179149102307, 0.30625420155821903, 1.288208276267654,
             1.3345939131355147, -1.4713774812532359]
    x = [0.0, 0.0, 1.1489955608128035, -9.641815125514444, 0.5975150739415511, 5.321795555035614,
         -5.467954909319733, 1.4687325444642383, 8.370366165425736, -2.810816575553474, 1.815504174076466,
         1.1769187743854737, -8.870585410511126, -0.14181686526567816, 7.93970270970086, 0.14183234352470464,
         -8.054425087325441, -3.809075460321715]
    want = [0., 0., 109.78978122898833, 32.2780625869145, -32.38838457188155, 33.25158538253831,
            -12.072197714155793, -6.87521436567669, 26.39304084900066, -2.6761855166330424, 1.8718030426410628,
            11.712151684565809, -23.362167910509047, -13.601765954771285, 21.606370954127627, 12.97052379948305,
            -12.477808805315766, -4.962089239867884]
    seg = _seg(2, 4, 10, 2, lcols=5, lcol0=2, lcoef=0, rcols=3, rcol0=2 + 16 - 3, rcoef=20)
    got = _block_apply(ctx, 18, [seg], left + right, interior, np.array(x), np.zeros(18), 0)
    _close(got, want)


def test_csr_random(ctx):
    """csr.t.cpp:52-95: 10 rows, 9 entries (rows 3-5 and 7-8 empty); apply
    adds into b."""
    w = [6.132558989050928, -0.4611523807581932, -2.874686661596037, 9.42084557206411, 0.2298026797436883,
         6.066446959605997, -7.70721485928825, -0.9885546582519957, -5.302176517574914]
    v = [1, 6, 0, 4, 6, 7, 8, 9, 0]
    u = [0, 2, 3, 4, 4, 4, 4, 8, 8, 8, 9]
    x = [-3.612622416000683, -1.1427601879942273, 0.8552565615441736, -4.755547850496647, -9.711932690401355,
         -7.232904766266888, -5.437050079342718, -2.4092054560513105, 3.557741982344453, 5.127028269794991]
    want = [-4.500735674823108, 10.385157472660014, -91.49461808255228, 0., 0., 0., -48.353395342996556, 0., 0.,
            19.15476174098357]
    ua, va, wa = (np.array(u, np.int64), np.array(v, np.int64), np.array(w))
    A = C.c_void_p()
    assert lib().shx_csr_create(ctx.p, 10, ua.ctypes.data_as(i64p), va.ctypes.data_as(i64p),
                                wa.ctypes.data_as(dp), C.byref(A)) == 0
    px, pb = _dev(ctx, x), _dev(ctx, np.zeros(10))
    assert lib().shx_csr_apply(A, px, pb) == 0
    lib().shx_sync(ctx.p)
    got = _host(ctx, pb, 10)
    _close(got, want)
    # always +=: a second apply doubles
    assert lib().shx_csr_apply(A, px, pb) == 0
    lib().shx_sync(ctx.p)
    _close(_host(ctx, pb, 10), 2 * np.array(want))
    lib().shx_free(ctx.p, px)
    lib().shx_free(ctx.p, pb)
    lib().shx_csr_destroy(A)
