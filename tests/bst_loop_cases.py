"""Arguments of the reference's BST test (test/test_bst.jl:72-85) and the
energy it pins (:87, atol 1e-6)."""
import numpy as np

H12_ARGS = dict(max_iter=20, nbody=4, H0="Hcmf", thresh_var=1e-2, thresh_foi=1e-3,
                thresh_pt=np.sqrt(1e-5), ci_conv=1e-5, do_pt=True, resolve_ss=True,
                tol_tucker=1e-4)
E_VAR_REF = -18.329455496968325
