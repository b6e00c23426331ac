"""CPU checks of the learned permutation allocation: the float64 reference
network (oracle/mlp_ref.py) against finite differences and closed forms, and
the allocation policy (budget levels, monotonicity, moments arithmetic)."""
import math

import numpy as np
import pytest

import mlp_ref


def test_mlp_ref_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp_ref.init_params(1).astype(np.float64) + 0.05 * rng.standard_normal(mlp_ref.NUM_PARAMS)
    x = rng.standard_normal((7, mlp_ref.IN))
    t = rng.standard_normal((7, mlp_ref.OUT))
    _, g = mlp_ref.loss_and_grad(p, x, t)
    # central differences on a spread of parameters from every tensor
    idx = np.concatenate([rng.choice(mlp_ref.NUM_PARAMS, 60, replace=False),
                          [0, 1023, 1024, 1087, 1088, 5183, 5184, 5439, 5440, 5507]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mlp_ref.loss_and_grad(p + e, x, t)
        lm, _ = mlp_ref.loss_and_grad(p - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), i


def test_mlp_ref_closed_forms():
    # gelu: gelu(0) = 0, gelu(x) - gelu(-x) = x, gelu'(0) = 1/2
    z = np.linspace(-4, 4, 33)
    np.testing.assert_allclose(mlp_ref.gelu(z) - mlp_ref.gelu(-z), z, atol=1e-12)
    assert mlp_ref.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp_ref.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    # zero weights: the output is the last bias for every input
    d = {n: np.zeros(s) for n, s in mlp_ref.SHAPES}
    d["b3"] = np.array([1.0, -2.0, 0.5, 3.0])
    y, _ = mlp_ref.forward(mlp_ref.pack(d), np.random.default_rng(1).standard_normal((5, 16)))
    np.testing.assert_array_equal(y, np.tile(d["b3"], (5, 1)))
    # a single linear path: W1 picks x0, W2 passes h1_0, W3 reads h2_0
    d = {n: np.zeros(s) for n, s in mlp_ref.SHAPES}
    d["W1"][0, 0] = 1.0
    d["W2"][0, 0] = 1.0
    d["W3"][2, 0] = 1.0
    x = np.zeros((1, 16))
    x[0, 0] = 1.3
    y, _ = mlp_ref.forward(mlp_ref.pack(d), x)
    g1 = 0.5 * 1.3 * (1 + math.erf(1.3 / math.sqrt(2)))
    g2 = 0.5 * g1 * (1 + math.erf(g1 / math.sqrt(2)))
    assert abs(y[0, 2] - g2) < 1e-12 and y[0, 0] == 0.0


def test_param_count_matches_header():
    import re
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "gsea_hip.h")).read()
    n = int(re.search(r"#define GSEA_MLP_NUM_PARAMS (\d+)", hdr).group(1))
    from cudagsea_amd import surrogate as sg
    assert n == mlp_ref.NUM_PARAMS == sg.NUM_PARAMS
    assert sg.init_params(0).size == n


def test_null_abs_stats_closed_form():
    from cudagsea_amd import surrogate as sg
    # symmetric tails with mean m and std s: mean |ES| = m, cv = s / m
    mom = np.array([[0.4, 0.1, 0.4, 0.1], [0.2, 0.0, 0.2, 0.0]])
    mean, cv = sg.null_abs_stats(mom)
    np.testing.assert_allclose(mean, [0.4, 0.2])
    np.testing.assert_allclose(cv, [0.25, 0.0], atol=1e-12)
    # asymmetric tails: mixture of two components
    mom = np.array([[0.3, 0.1, 0.5, 0.2]])
    mean, cv = sg.null_abs_stats(mom)
    m2 = 0.5 * (0.01 + 0.09 + 0.04 + 0.25)
    assert abs(mean[0] - 0.4) < 1e-12
    assert abs(cv[0] - math.sqrt(m2 - 0.16) / 0.4) < 1e-12


def test_budget_policy():
    from cudagsea_amd import surrogate as sg
    P = 16384
    G = 200
    rng = np.random.default_rng(3)
    mom = np.tile([0.3, 0.09, 0.3, 0.09], (G, 1))
    es0 = np.linspace(0.0, 0.9, G) * np.where(rng.random(G) < 0.5, -1, 1)
    b = sg.budgets(mom, es0, P, nes_tol=0.02, z=4.5, margin=0.0, min_budget=512)
    levels = {P // 2 ** i for i in range(6)}
    assert set(np.unique(b)) <= levels
    assert (P % b == 0).all() and b.min() >= 512 and b.max() <= P
    # non-decreasing in |NES|
    o = np.argsort(np.abs(es0), kind="stable")
    assert (np.diff(b[o]) >= 0).all()
    # the rule itself: z*cv*|NES|/sqrt(B) <= tol at the chosen level, and
    # fails one level below (unless at the floor)
    mean, cv = sg.null_abs_stats(mom)
    nes = np.abs(es0) / mean
    se = 4.5 * cv * nes / np.sqrt(b)
    assert (se[b < P] <= 0.02 + 1e-12).all()
    below = (b > 512)
    se_half = 4.5 * cv * nes / np.sqrt(b // 2)
    assert (se_half[below] > 0.02).all()
    # a margin never lowers a budget
    b2 = sg.budgets(mom, es0, P, margin=0.2)
    assert (b2 >= b).all()
    # non-finite predictions fall back to the full count
    bad = mom.copy()
    bad[0] = np.nan
    assert sg.budgets(bad, es0, P)[0] == P
