"""The reference's own data (test_PDE/Data_Gray_Scott, read through
vcphip_mat_read) as operands of the GPU products: dd products of the
Gray-Scott solution coefficients and interval products of point intervals
made from them, bit-identical to the oracle."""
import os

import numpy as np
import pytest

import oracle_ffi as O
import vcp_amd as V

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vcp_data")


@pytest.mark.parametrize("name", ["uh_2D_solution.matrix_kvdd", "uh_Graf_solution1_eps55.matrix_kvdd"])
def test_dd_products_of_reference_data(ctx, name):
    U, t = V.mat_read(os.path.join(DATA, name))
    assert t == V.MAT_DD
    n = U.shape[0]
    G = ctx.ddgemm_tn(U)                  # U^T U
    assert np.array_equal(G, O.ozaki_ddgemm(U, None, O.dd_default_nmod(n)))
    Ut = O.pairs(U.shape[1], n)
    Ut[...] = U.transpose(1, 0, 2)
    P = ctx.ddgemm(U, Ut)                 # U U^T (k = number of columns)
    assert np.array_equal(P, O.ozaki_ddgemm(U, Ut, O.dd_default_nmod(U.shape[1])))


def test_interval_products_of_reference_data(ctx):
    D, t = V.mat_read(os.path.join(DATA, "uh_2D_solution.matrix_d"))
    assert t == V.MAT_DOUBLE
    IA = O.pairs(*D.shape)
    IA[..., 0] = D
    IA[..., 1] = np.nextafter(D, np.inf)
    IT = O.pairs(D.shape[1], D.shape[0])
    IT[...] = IA.transpose(1, 0, 2)
    for form, ref in ((V.FORM_PLAIN, O.pidblas_mulmm(IT, IA)), (V.FORM_OZAKI, O.ozaki_igemm(IT, IA, 16))):
        C = ctx.igemm(IT, IA, form)
        assert np.array_equal(C, ref)
