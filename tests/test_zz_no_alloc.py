"""The allocation-free soft-contact entry point (vbd_collide_generate_noalloc) against vbd_collide_generate.

Both are called through the C ABI on the same positions; the records, their order and the
contact count must be identical.  Two calls in a row on one device share the static scratch:
a 256x256 cloth pressed into the sphere and the ground (131,072 pairs, 128 tiles), then a
32x32 cloth resting on the ground (1,024 pairs, one tile), so the second call runs with
fewer tiles over scratch the first call left filled.  A third case overflows the buffer
(soft_contact_max below the hit count).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("soft_contact_particle", "soft_contact_indices", "soft_contact_barycentric", "soft_contact_shape",
          "soft_contact_body_pos", "soft_contact_body_vel", "soft_contact_normal")


def _fill_garbage(c):
    import torch

    for name in FIELDS:
        t = getattr(c, name)
        if t.dtype == torch.int32:
            t.fill_(-7)
        else:
            t.fill_(float("nan"))
    c.soft_contact_count.fill_(-7)


def _compare(model, q, margin, capacity=None):
    import torch

    from newton_amd import _native
    from newton_amd.sim.collide import CollisionPipeline, contacts_out
    from newton_amd.sim.model import Contacts, State

    pipe = CollisionPipeline(model, soft_contact_margin=margin)
    cap = pipe.soft_contact_max if capacity is None else capacity
    ref, got = Contacts(cap, model.device), Contacts(cap, model.device)
    _fill_garbage(got)
    st = State()
    st.particle_q = q
    pipe.collide(st, ref)
    stream = torch.cuda.current_stream(model.device).cuda_stream
    rc = _native.hip_lib().vbd_collide_generate_noalloc(ctypes.byref(pipe.desc), stream, _native.ptr(q), margin,
                                                        ctypes.byref(contacts_out(got)))
    _native.check(rc, "vbd_collide_generate_noalloc")
    torch.cuda.synchronize()
    n_ref = int(ref.soft_contact_count.item())
    n_got = int(got.soft_contact_count.item())
    assert n_got == n_ref
    k = min(n_ref, cap)
    for name in FIELDS:
        a = getattr(ref, name)[:k].cpu().numpy()
        b = getattr(got, name)[:k].cpu().numpy()
        assert np.array_equal(a.view(np.int32), b.view(np.int32)), name
    return n_ref, pipe.soft_rigid_contact_pair_count


def test_noalloc_collide_matches_allocating_entry_point(cuda_device):
    from newton_amd import scenes

    # 1) 256x256 cloth (2 m) lowered to y = 0.3: its centre lies inside the r = 0.4 sphere
    model, state, _ = scenes.cloth_drop_on_sphere(cuda_device)
    q = state.particle_q.clone()
    q[:, 1] -= 1.2
    n1, p1 = _compare(model, q, 0.01)
    assert p1 == 2 * model.particle_count and n1 > 1000
    print(f"call 1: {p1} pairs, {n1} contacts")
    # 2) a smaller input on the same device: 32x32 cloth resting on the ground
    model2, state2, _ = scenes.cloth_hanging(cuda_device)
    q2 = state2.particle_q.clone()
    q2[:, 1] = 0.05
    n2, p2 = _compare(model2, q2, 0.01)
    assert p2 < p1 and n2 == int(((model2.particle_flags & 1) != 0).sum().item())
    print(f"call 2: {p2} pairs, {n2} contacts")
    # 3) overflow: the count is the number of hits, records stop at the capacity
    n3, _ = _compare(model, q, 0.01, capacity=500)
    assert n3 == n1


def test_noalloc_collide_rejects_oversize(cuda_device):
    from newton_amd import _native

    d = _native.VbdCollideDesc()
    d.pair_count = 16 * 1024 * 1024 + 1
    out = _native.VbdContactsOut()
    cnt = ctypes.c_int32(0)
    out.soft_contact_count = ctypes.addressof(cnt)  # never written: the call fails on the size check
    rc = _native.hip_lib().vbd_collide_generate_noalloc(ctypes.byref(d), None, ctypes.c_void_p(8), 0.01,
                                                        ctypes.byref(out))
    assert rc == -1
    msg = _native.hip_lib().vbd_last_error().decode()
    assert "exceeds the maximum" in msg
