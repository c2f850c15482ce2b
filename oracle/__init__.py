'''
Oracle: a CPU, float64 (and longdouble for the Legendre recurrence)
restatement of optweight's conjugate-gradient Wiener-filter path.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker (or, in ``bench.py``, as the timed CPU
baseline). The product (``optweight_amd``) never imports it and fails loudly
when its HIP library is missing.

Modules and the reference code each one restates (paths relative to the
optweight repository):

  layout     alm indexing, ``alm_c_utils.c:3-10``, ``hp.Alm.getlmax``
  geometry   ``map_utils.MapInfo`` 19-297, ``get_gauss_minfo`` 914-941,
             ``get_equal_area_minfo`` 943-992, ``get_fullsky_geometry``
             349-413, ``minfo2lmax`` 1795-1811, ``get_mtype`` 299-347
  legendre   normalised (spin-weighted) associated Legendre functions used by
             ducc0's synthesis (called at ``sht.py:78-93,159-174``)
  sht        ``sht.alm2map`` 95-178 / ``sht.map2alm`` 10-93 (ducc0 synthesis
             and adjoint_synthesis), ``default_spin`` 180-215
  ops        ``alm_c_utils.lmul`` (``alm_c_utils.c:12-217``),
             ``trunc_alm``/``wlm2alm``, ``map_c_utils.apply_ringweight``
             (``map_c_utils.c:7-37``), ``mat_utils.matvec/matpow/full_matrix``,
             ``alm_utils.contract_almxblm`` 9-44, ``unit_var_alm``/``rand_alm``
             247-338, ``map_utils.rand_map_pix`` 1084-1122,
             ``get_isotropic_ivar`` 1172-1218, ``inv_qweight_map`` 1124-1170,
             ``noise_utils._band_limit_gauss_beam`` 401-421,
             ``multigrid.lowpass_filter`` 449-471
  cg         ``pixell.utils.CG`` as used by ``solvers.py:127-141``
  wavelet    ``alm_utils.alm2wlm_axisym``/``wlm2alm_axisym`` 46-155,
             ``operators.WavMatVecAlm`` 705-815, ``EllWavEllMatVecAlm``
             645-703, ``map_utils.gauss2gauss`` 672-717, ``get_ivar_ell``
             1220-1266, ``preconditioners.PseudoInvPreconditionerWav``
             251-395, ``get_itau_ell_harm`` 916-948, ``solvers.CGWiener``
             697-1139
  wiener     ``solvers.CGWienerMap`` 8-456, ``preconditioners``
             (Harmonic 8-124, PseudoInv 126-249, Masked 558-650),
             ``multigrid`` 13-447, ``operators`` (PixEllPixMatVecMap 991-1102,
             PixMatVecAlm 562-643, add_operators 1458-1489)

How each part is pinned (none is "parity unpinned" unless listed here):
  * lmul / trunc_alm / wlm2alm / apply_ringweight: against the reference's
    own C sources compiled into ``oracle/_ref`` (recipe ``oracle/Makefile``)
    and the reference tests' literal arrays (``test_alm_c_utils.py:10-59,
    387-441``).
  * contract_almxblm: ``test_alm_utils.py:411-448`` (-40, -120, -240).
  * geometry: ``scipy.special.roots_legendre`` and sum(w*nphi) = 4 pi.
  * Legendre functions: an exact Wigner-d sum (mpmath, 40 digits) for
    l <= 12, the textbook values Y_00, Y_11, 2Y_20, and the independent
    HEALPix closed form of F1/F2 (Zaldarriaga & Seljak 1997) for spin 2.
  * SHT: brute-force synthesis with the exact functions, round trips
    Y^T W Y = I on GL grids, the dot-product adjoint identity.
  * CG / Wiener / preconditioners: closed forms of ``test_solvers.py:45-175``
    and full-sky white-noise exact inverses (CG converges in one step).
  * Gaussian beam: ``test_noise_utils.py:120-135`` golden values.
  * multigrid: ``test_multigrid.py`` / ``test_preconditioners.py`` properties.
  * wavelet noise model and harmonic-space solver: per-multipole closed
    forms for per-scale inverse covariance proportional to the quadrature
    weights (WavMatVecAlm, its weighted inverse, get_ivar_ell, one-step
    exact CG, the wavelet pseudo-inverse), a planted missing-weight check.
'''
