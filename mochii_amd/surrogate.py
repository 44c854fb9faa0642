"""Learned warm start of the per-leaf balance solve: data collection from the
library's own models, initialisation, training on the GPU, storage.

The network (8 -> 64 -> 64 -> 64 -> 4, GELU) maps a leaf's tallied rates to
a first guess of (log10 x_HI, log10 x_HeI, log10 x_HeII, log10 T_e); the
thermal update's safeguarded Newton solve starts from it (MC_WARM_MLP) and
judges every leaf.  Training data are cell states the transport + thermal
iteration itself produces: the features before an update and the
reference-equivalent bisection solve's result after it.  The reference
program has no learned component; this module is the build's own.
"""
import numpy as np

from .lib import MC_MLP_NPARAM, MC_WARM_NONE

NI, NH, NO = 8, 64, 4
_SIZES = [NI * NH, NH, NH * NH, NH, NH * NH, NH, NH * NO, NO, NI, NI, NO, NO]
_OFFS = np.concatenate([[0], np.cumsum(_SIZES)])
assert _OFFS[-1] == MC_MLP_NPARAM


def targets_of(state):
    """Training targets from a leaf state: log10 x_HI, x_HeI, x_HeII, T_e."""
    f = lambda a: np.log10(np.clip(a, 1e-12, None))
    return np.stack([f(state["xHI"]), f(state["xHeI"]), f(state["xHeII"]),
                     np.log10(state["Te"])], axis=1).astype(np.float32)


def collect(model, iterations, every=1, log=None):
    """Run `iterations` transport + thermal iterations of a set-up Model with
    the plain solve (relaxation 1 is what the targets assume) and return the
    (features, targets) rows of leaves with gas, every `every`-th iteration."""
    model.set_warm_start(MC_WARM_NONE)
    Xs, Ys = [], []
    for it in range(iterations):
        model.ionizing_field_and_rates()
        X = model.mlp_features()
        m = model.gas_thermal_update()
        if it % every == 0:
            st = model.state("nH", "xHI", "xHeI", "xHeII", "Te")
            keep = st["nH"] > 0
            Xs.append(X[keep])
            Ys.append(targets_of(st)[keep])
        if log:
            log(f"  collect iteration {it}: dx_vol {m[2]:.3e} dte_vol {m[3]:.3e}")
    return np.concatenate(Xs), np.concatenate(Ys)


def init_params(X, Y, seed=0):
    """He-normal weights, zero biases, standardisation from the data."""
    rng = np.random.default_rng(seed)
    P = np.zeros(MC_MLP_NPARAM, dtype=np.float64)
    fan = [NI, NH, NH, NH]
    for l in range(4):
        w0 = _OFFS[2 * l]
        n = _SIZES[2 * l]
        P[w0:w0 + n] = rng.normal(0.0, np.sqrt(2.0 / fan[l]), n)
    mu_x, sd_x = X.mean(axis=0), X.std(axis=0)
    mu_y, sd_y = Y.mean(axis=0), Y.std(axis=0)
    P[_OFFS[8]:_OFFS[9]] = mu_x
    P[_OFFS[9]:_OFFS[10]] = np.maximum(sd_x, 1e-3)
    P[_OFFS[10]:_OFFS[11]] = mu_y
    P[_OFFS[11]:_OFFS[12]] = np.maximum(sd_y, 1e-3)
    return P.astype(np.float32)


def balance(X, Y, max_rows=400_000, pinned_frac=0.15, seed=0):
    """Subsample: every leaf off the temperature floor, and a fraction of the
    (many, identical) leaves pinned at te_min, to at most max_rows rows."""
    rng = np.random.default_rng(seed)
    floor = Y[:, 3] <= Y[:, 3].min() + 1e-6
    hot = np.flatnonzero(~floor)
    cold = np.flatnonzero(floor)
    ncold = min(cold.size, int(pinned_frac * max(hot.size, 1)) + 1)
    idx = np.concatenate([hot, rng.choice(cold, ncold, replace=False) if ncold else cold[:0]])
    if idx.size > max_rows:
        idx = rng.choice(idx, max_rows, replace=False)
    rng.shuffle(idx)
    return X[idx], Y[idx]


def train(model, X, Y, params=None, schedule=((3000, 2e-3), (2000, 4e-4)), batch=2048,
          seed=1, log=None):
    """Adam on the GPU through the library's training kernels; returns the
    trained parameter blob and the loss history."""
    P = init_params(X, Y, seed) if params is None else np.asarray(params, dtype=np.float32)
    model.mlp_set(P)
    hist = []
    for k, (steps, lr) in enumerate(schedule):
        h = model.mlp_train(X, Y, steps, batch=batch, lr=lr, seed=seed + k)
        hist.append(h)
        if log:
            log(f"  train phase {k}: {steps} steps at lr {lr:g}: loss {h[:50].mean():.3e} -> "
                f"{h[-50:].mean():.3e}")
    return model.mlp_get(), np.concatenate(hist) if hist else np.zeros(0)


def save(path, params):
    np.save(path, np.asarray(params, dtype=np.float32), allow_pickle=False)


def load(path):
    P = np.load(path, allow_pickle=False)
    if P.size != MC_MLP_NPARAM:
        raise ValueError(f"{path}: expected {MC_MLP_NPARAM} parameters, found {P.size}")
    return P.astype(np.float32)
