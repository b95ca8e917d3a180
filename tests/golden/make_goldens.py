"""Write the golden vectors of tests/golden/ from the reference's own C kernels.

Runs in the development container only (it needs /root/reference and the
library oracle/_ref/libctools_ref.so compiled from
/root/reference/fem_c/ctools.c by ``make -C oracle``).  Inputs are tiny
synthetic cases built with this project's mesher and closed-form tables; the
outputs are what the reference's compute_K_dK_Hcurl_red / compute_K_dK_Hcurl /
compute_B_Hcurl / compute_B_line_segs / eval_pchip_interpolator(_derivative)
return for them.  Only data is written (npz / json).

    OMP_NUM_THREADS=1 python tests/golden/make_goldens.py
"""
import ctypes
import json
import os
import sys

import numpy as np

os.environ.setdefault('OMP_NUM_THREADS', '1')
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from oracle import oracle as orc  # noqa: E402
from snoopy_amd import finite_element as fe  # noqa: E402
from snoopy_amd.magnets import Dipole  # noqa: E402
from snoopy_amd.materials import Reluctance, model_tables  # noqa: E402

REF_MAT = '/root/reference/files/materials/aisi1010.json'
P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731


def bh_fixture():
    d = json.load(open(REF_MAT))
    out = {'B(T)': d['BH_data']['B(T)'], 'H(A/m)': d['BH_data']['H(A/m)'],
           'source': 'files/materials/aisi1010.json (BH_data)'}
    json.dump(out, open(os.path.join(HERE, 'aisi1010_bh.json'), 'w'))
    return Reluctance(B_data=out['B(T)'], H_data=out['H(A/m)'])


def small_case(seed=0, ne_iron=16, ne_air=16):
    dp = Dipole(h=0.25)
    m = dp.mesh
    iron = np.nonzero(m.tags == 1)[0][:ne_iron]
    air = np.nonzero(m.tags == 2)[0][:ne_air]
    sel = np.concatenate([iron, air])
    # compact node and edge numbering
    nodes_used, cells = np.unique(m.cells[sel], return_inverse=True)
    cells = cells.reshape(-1, 4).astype(np.int32)
    edges_used, gids = np.unique(m.glob_ids[sel], return_inverse=True)
    gids = gids.reshape(-1, 6).astype(np.int32)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(len(edges_used)) * 2e-2
    return dict(nodes=m.nodes[nodes_used], cells=cells, gids=gids, orient=m.orient[sel],
                n_iron=len(iron), x=x, num_dofs=len(edges_used))


def ref_elements(R, c, q_order, rel, Bs, part):
    w, q = fe.quadrature_rule(q_order)
    Q = len(w)
    curls = np.ascontiguousarray(fe.curl_table(Q))
    dphi = np.ascontiguousarray(fe.lagrange_basis_derivative(q).reshape(-1))
    ts, n_int, cf, kn = model_tables(rel)
    cf = np.ascontiguousarray(cf)
    kn = np.ascontiguousarray(kn)
    if part == 'iron':
        sl = slice(0, c['n_iron'])
    else:
        sl = slice(c['n_iron'], None)
    cells = np.ascontiguousarray(c['cells'][sl].reshape(-1), dtype=np.int32)
    gids = np.ascontiguousarray(c['gids'][sl].reshape(-1), dtype=np.int32)
    orient = np.ascontiguousarray(c['orient'][sl], dtype=np.int32)
    ne = len(orient)
    nt = 36 * ne
    rk, ck, vk = np.zeros(nt, np.int32), np.zeros(nt, np.int32), np.zeros(nt)
    rd, cd, vd = np.zeros(nt, np.int32), np.zeros(nt, np.int32), np.zeros(nt)
    nodes = np.ascontiguousarray(c['nodes'].reshape(-1))
    x = np.ascontiguousarray(c['x'])
    tK = R.make_triplet_list(nt, P(rk), P(ck), P(vk))
    tD = R.make_triplet_list(nt, P(rd), P(cd), P(vd))
    rl = R.make_reluctance(ts, n_int, P(cf), P(kn))
    msh = R.make_mesh(len(c['nodes']), ne, 4, P(nodes), P(cells))
    qq, ww = np.ascontiguousarray(q), np.ascontiguousarray(w)
    qr = R.make_quadrature_rule(Q, P(qq), P(ww))
    rhs = np.zeros(c['num_dofs'])
    if Bs is not None:
        bs = np.ascontiguousarray(Bs)
        R.compute_K_dK_Hcurl_red(tK, tD, P(rhs), P(gids), rl, msh, qr, P(curls), P(dphi),
                                 P(orient), P(x), P(bs))
    else:
        R.compute_K_dK_Hcurl(tK, tD, P(gids), rl, msh, qr, P(curls), P(dphi), P(orient), P(x))
    return dict(rows=rk, cols=ck, vK=vk, vdK=vd, rhs=rhs, w=w, q=q)


def main():
    R = orc.ref_lib()
    if R is None:
        raise SystemExit('oracle/_ref/libctools_ref.so missing: run make -C oracle')
    rel = bh_fixture()
    c = small_case()
    out = {k: v for k, v in c.items()}
    rng = np.random.default_rng(1)
    for order in (2, 3):
        Q = len(fe.quadrature_rule(order)[0])
        Bs = rng.standard_normal(c['n_iron'] * Q * 3) * 0.8
        g = ref_elements(R, c, order, rel, Bs, 'iron')
        for k in ('rows', 'cols', 'vK', 'vdK', 'rhs'):
            out['iron_q%d_%s' % (order, k)] = g[k]
        out['iron_q%d_Bs' % order] = Bs
    from snoopy_amd.materials import ConstantReluctance, NU0
    g = ref_elements(R, c, 2, ConstantReluctance(NU0), None, 'air')
    for k in ('rows', 'cols', 'vK', 'vdK'):
        out['air_q2_%s' % k] = g[k]
    # B at quadrature points (compute_B_Hcurl, ctools.c:1500-1563)
    w, q = fe.quadrature_rule(2)
    Q = len(w)
    cells = np.ascontiguousarray(c['cells'].reshape(-1), dtype=np.int32)
    gids = np.ascontiguousarray(c['gids'].reshape(-1), dtype=np.int32)
    orient = np.ascontiguousarray(c['orient'], dtype=np.int32)
    ne = len(orient)
    pts = np.zeros(ne * Q * 3)
    B = np.zeros(ne * Q * 3)
    nodes = np.ascontiguousarray(c['nodes'].reshape(-1))
    msh = R.make_mesh(len(c['nodes']), ne, 4, P(nodes), P(cells))
    qq, ww = np.ascontiguousarray(q), np.ascontiguousarray(w)
    qr = R.make_quadrature_rule(Q, P(qq), P(ww))
    curls = np.ascontiguousarray(fe.curl_table(Q))
    phi = np.ascontiguousarray(fe.lagrange_basis(q).reshape(-1))
    dphi = np.ascontiguousarray(fe.lagrange_basis_derivative(q).reshape(-1))
    x = np.ascontiguousarray(c['x'])
    R.compute_B_Hcurl(P(pts), P(B), 6, P(gids), msh, qr, P(curls), P(phi), P(dphi), P(orient), P(x))
    out['bhcurl_points'] = pts.reshape(-1, 3)
    out['bhcurl_B'] = B.reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, 'elements.npz'), **out)

    # PCHIP at knots, between knots, beyond the last knot
    ts, n_int, cf, kn = model_tables(rel)
    cf, kn = np.ascontiguousarray(cf), np.ascontiguousarray(kn)
    pts_b = np.concatenate([kn[:40], 0.5 * (kn[:-1] + kn[1:])[::4], [kn[-1], kn[-1] * 1.5, 150.0]])
    nu = np.array([R.eval_pchip_interpolator(n_int, P(cf), P(kn), float(b)) for b in pts_b])
    dnu = np.array([R.eval_pchip_interpolator_derivative(n_int, P(cf), P(kn), float(b))
                    for b in pts_b])
    np.savez_compressed(os.path.join(HERE, 'pchip.npz'), B=pts_b, nu=nu, dnu=dnu, c=cf, x=kn)

    # Biot-Savart: racetrack segments plus hand-made segments; targets far,
    # inside the core and near an end point (ctools.c:1869-1902 branches)
    from snoopy_amd.racetrack_coil import RacetrackCoil
    kp = np.array([[-0.1, 0.2], [-0.05, 0.25], [0.05, 0.25], [0.1, 0.2], [0.1, -0.2],
                   [0.05, -0.25], [-0.05, -0.25], [-0.1, -0.2]])
    coil = RacetrackCoil(kp, np.array([-0.01, 0.0, 0.01]), 0.004, 123.0)
    segs = coil.get_segments()
    segs = np.concatenate([segs, [[0.3, 0.0, -0.1, 0.3, 0.0, 0.1], [0.0, 0.4, 0.0, 0.2, 0.4, 0.0]]])
    tar = np.concatenate([
        rng.uniform(-0.4, 0.4, size=(60, 3)),
        np.array([[0.3 + 0.001, 0.002, 0.0], [0.3, 0.0005, 0.05], [0.3 + 0.002, 0.0, 0.1 + 0.001],
                  [0.1 + 0.002, 0.0, 0.0], [-0.1, 0.003, 0.1], [0.0, 0.4 + 0.001, 0.0],
                  [0.2 + 0.0015, 0.4, 0.0]])])
    src = np.ascontiguousarray(segs.reshape(-1))
    tt = np.ascontiguousarray(tar.reshape(-1))
    Bb = np.zeros(len(tar) * 3)
    R.compute_B_line_segs(P(Bb), P(src), P(tt), 123.0, 0.004, len(segs), len(tar))
    np.savez_compressed(os.path.join(HERE, 'biot_savart.npz'), segs=segs, tar=tar,
                        current=123.0, rad=0.004, B=Bb.reshape(-1, 3))
    print('goldens written to', HERE)


if __name__ == '__main__':
    main()
