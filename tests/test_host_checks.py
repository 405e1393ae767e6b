"""Host-side argument checks of the torch-facing calls (no GPU): undersized,
mis-typed or mis-shaped buffers of the batched checkpointed forward, the
batched adjoint and the single-configuration adjoint raise before any launch."""
import pytest
import torch

from emtrace import api, autograd
from emtrace.api import BatchLayout, check_batch_buffers


def _layout():
    return BatchLayout.of([5, 3], [4, 2], [6, 3])   # 8 charges, 4*6 + 2*3 = 30 line elements


def _ok(layout):
    return dict(charges=torch.zeros(layout.total_charges, 4),
                lines=torch.zeros(3, layout.total_line_elems),
                steps=torch.zeros(6, dtype=torch.int32), ckpt=torch.zeros(100), ckpt_floats=100,
                grad_lines=torch.zeros(3, layout.total_line_elems),
                grad_charges=torch.zeros(layout.total_charges, 4))


def test_batch_buffers_accepted():
    L = _layout()
    check_batch_buffers(L, **_ok(L), on_device=False)


@pytest.mark.parametrize("name,bad,exc", [
    ("charges", torch.zeros(7, 4), ValueError),
    ("charges", torch.zeros(8, 3), ValueError),
    ("charges", torch.zeros(8, 4, dtype=torch.float64), TypeError),
    ("lines", torch.zeros(3, 29), ValueError),
    ("lines", torch.zeros(3, 60)[:, ::2], ValueError),            # not contiguous
    ("steps", torch.zeros(5, dtype=torch.int32), ValueError),
    ("steps", torch.zeros(6, dtype=torch.int64), TypeError),
    ("ckpt", torch.zeros(99), ValueError),
    ("grad_lines", torch.zeros(3, 29), ValueError),
    ("grad_lines", torch.zeros(3, 30, dtype=torch.float64), TypeError),
    ("grad_charges", torch.zeros(7, 4), ValueError),
    ("grad_charges", torch.zeros(16, 2), ValueError),
])
def test_batch_buffers_refused(name, bad, exc):
    L = _layout()
    args = _ok(L)
    args[name] = bad
    with pytest.raises(exc):
        check_batch_buffers(L, **args, on_device=False)


def test_batch_buffers_on_device_refuses_host_tensors():
    L = _layout()
    with pytest.raises(TypeError):
        check_batch_buffers(L, **_ok(L))


def _adj(n=10, length=12, every=5):
    need = api.ckpt_size(n, length, every)
    return dict(charges=torch.zeros(4, 4), rows=torch.zeros(3, length, n),
                ckpt=torch.zeros(need), ckpt_every=every, grad_rows=torch.zeros(3, length, n))


def test_adjoint_buffers():
    autograd.check_adjoint_buffers(**_adj(), on_device=False)
    for key, bad in (("ckpt", None), ("grad_rows", torch.zeros(3, 11, 10)),
                     ("charges", torch.zeros(4, 3)), ("rows", torch.zeros(2, 12, 10))):
        a = _adj()
        a[key] = torch.zeros(a["ckpt"].numel() - 1) if bad is None else bad
        with pytest.raises(ValueError):
            autograd.check_adjoint_buffers(**a, on_device=False)
    a = _adj()
    a["grad_rows"] = torch.zeros(3, 12, 10, dtype=torch.float64)
    with pytest.raises(TypeError):
        autograd.check_adjoint_buffers(**a, on_device=False)
    a = _adj()
    a["ckpt_every"] = 0
    with pytest.raises(ValueError):
        autograd.check_adjoint_buffers(**a, on_device=False)
