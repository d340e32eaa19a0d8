"""Learned sweep cost model (include/q1tsim_gvs.h, q1t_cm_*).

A 20 -> 64 -> 64 -> 1 tanh MLP whose forward, backward and Adam kernels are
hand-written HIP (csrc/costmodel.hip).  It predicts log(microseconds per
1024 tiles) of a sweep program from the features csrc/plan.cpp computes
(sweep_features); the fused complex64 planner uses it to rank equivalent
programs.  Training data are sweeps the simulator times itself: a state with
recording on logs every sweep's features and HIP-event time.
"""
import ctypes
import json
import os

import numpy as np

from ._lib import check, lib

NFEAT = 20
NPARAMS = 5569
# the model trained on MI355X (gfx950) by train_costmodel.py
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel_gfx950.npz")


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class CostModel(object):
    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        check(lib().q1t_cm_new(int(seed), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().q1t_cm_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(NPARAMS, dtype=np.float32)
        check(lib().q1t_cm_get_params(self.h, _fptr(out), NPARAMS))
        return out

    @params.setter
    def params(self, p):
        p = _f32(p)
        check(lib().q1t_cm_set_params(self.h, _fptr(p), len(p)))

    def set_norm(self, mu, sigma):
        mu, sigma = _f32(mu), _f32(sigma)
        check(lib().q1t_cm_set_norm(self.h, _fptr(mu), _fptr(sigma), NFEAT))

    def get_norm(self):
        mu = np.zeros(NFEAT, dtype=np.float32)
        sg = np.zeros(NFEAT, dtype=np.float32)
        check(lib().q1t_cm_get_norm(self.h, _fptr(mu), _fptr(sg), NFEAT))
        return mu, sg

    def predict(self, x):
        x = _f32(x).reshape(-1, NFEAT)
        y = np.zeros(len(x), dtype=np.float32)
        check(lib().q1t_cm_predict(self.h, _fptr(x), len(x), _fptr(y)))
        return y

    def loss_grad(self, x, t):
        x = _f32(x).reshape(-1, NFEAT)
        t = _f32(t)
        g = np.zeros(NPARAMS, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        check(lib().q1t_cm_loss_grad(self.h, _fptr(x), _fptr(t), len(x), _fptr(g), NPARAMS, ctypes.byref(loss)))
        return loss.value, g

    def adam_step(self, x, t, lr=1e-3):
        x = _f32(x).reshape(-1, NFEAT)
        t = _f32(t)
        loss = ctypes.c_double(0.0)
        check(lib().q1t_cm_adam_step(self.h, _fptr(x), _fptr(t), len(x), float(lr), ctypes.byref(loss)))
        return loss.value

    def train(self, x, t, epochs=30, batch=256, lr=1e-3, seed=1):
        x = _f32(x).reshape(-1, NFEAT)
        t = _f32(t)
        loss = ctypes.c_double(0.0)
        check(lib().q1t_cm_train(self.h, _fptr(x), _fptr(t), len(x), int(epochs), int(batch), float(lr), int(seed),
                                 ctypes.byref(loss)))
        return loss.value

    def set_trained(self, on=True):
        check(lib().q1t_cm_set_trained(self.h, 1 if on else 0))


def targets(tiles, ms):
    """Training targets: log of microseconds per 1024 tiles."""
    return np.log(np.asarray(ms, dtype=np.float64) * 1e3 / (np.asarray(tiles, dtype=np.float64) / 1024.0))


def fit(cm, feat, tiles, ms, epochs=60, batch=256, lr=2e-3, seed=1, min_steps=600):
    """Standardise the features (constant ones get sigma 1), then train on
    log(us per 1024 tiles) for `epochs` passes, raised so that there are at
    least `min_steps` Adam steps however few the samples.  Returns the final
    mean squared error."""
    feat = np.asarray(feat, dtype=np.float64)
    steps_per_epoch = max(1, -(-len(feat) // batch))
    epochs = max(epochs, -(-min_steps // steps_per_epoch))
    mu = feat.mean(0)
    sg = feat.std(0)
    sg[sg < 1e-6] = 1.0
    cm.set_norm(mu, sg)
    return cm.train(feat, targets(tiles, ms), epochs=epochs, batch=batch, lr=lr, seed=seed)


def record(nr_bits, circuits, fusion_k=4, seed=1, reps=1, geometry=None, jit=False):
    """Run each circuit (a list of (name, qubits, params)) on a complex64
    state with the planner exploring its candidates, and return the recorded
    sweeps (features, tiles, ms).  `jit`: sweeps run by kernels specialised
    to their programs (each new program is compiled before its flush runs;
    the timings are the kernels' own HIP-event times)."""
    from .gvs import GpuVectorState, gate_matrix
    s = GpuVectorState(nr_bits, 1, "c64")
    s.set_plan(fusion_k, True)
    if jit:
        s.set_jit(True)
    if geometry is not None:
        s.set_geometry(*geometry)
    s.set_cost_model(None, 2, seed)
    s.set_recording(True)
    feats, tiles, ms = [], [], []
    for _ in range(reps):
        for ops in circuits:
            for g, b, p in ops:
                s.apply_gate(gate_matrix(g, p), b, g)
            s.flush()
            f, t, m = s.take_samples()
            feats.append(f)
            tiles.append(t)
            ms.append(m)
    s.close()
    return np.concatenate(feats), np.concatenate(tiles), np.concatenate(ms)


def errors(cm, feat, tiles, ms):
    """Mean squared error of the predicted log time, and the mean and 90th
    percentile of |predicted / measured - 1| in time."""
    if len(feat) == 0:
        return {"samples": 0}
    y = cm.predict(feat).astype(np.float64)
    t = targets(tiles, ms)
    rel = np.abs(np.exp(y - t) - 1.0)
    return {"samples": int(len(feat)), "mse_log_time": float(np.mean((y - t) ** 2)),
            "mean_rel_err": float(rel.mean()), "p90_rel_err": float(np.quantile(rel, 0.9))}


def save(cm, path, meta=None):
    """Write the parameters, the feature normalisation and a JSON metadata
    string (numpy .npz, no pickled objects)."""
    mu, sg = cm.get_norm()
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez(path, params=cm.params, mu=mu, sigma=sg, nfeat=np.int32(NFEAT),
             meta=np.array(json.dumps(meta or {})))


def load(path=None, seed=1):
    """The model stored by save() (default: the shipped MI355X model), marked
    trained; raises when the file is missing or was written for another
    feature count."""
    path = path or DEFAULT_MODEL
    with np.load(path, allow_pickle=False) as z:
        if int(z["nfeat"]) != NFEAT or z["params"].shape != (NPARAMS,):
            raise ValueError("%s: model of %d features / %d parameters, this library has %d / %d"
                             % (path, int(z["nfeat"]), z["params"].shape[0], NFEAT, NPARAMS))
        cm = CostModel(seed)
        cm.params = z["params"]
        cm.set_norm(z["mu"], z["sigma"])
        meta = json.loads(str(z["meta"]))
    cm.set_trained(True)
    cm.meta = meta
    return cm
