"""The hand-written streaming probes (spme_stream_probe, a measurement aid
with no counterpart in the reference) compute what they claim, so that the
rates the benchmark quotes are rates of real work: copy gives the source,
fill the value, the in-place scale value^(reps + 1) times the input (one
untimed launch, then reps), and the read probe's per-block sums add up to
the buffer's sum (small integers, so every FP32 sum is exact)."""
import pytest

from spme_amd.lib import SpmeError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.mark.parametrize("nontemporal", [False, True])
@pytest.mark.parametrize("blocks", [0, 8, 1000])
def test_stream_probes_compute_what_they_claim(torch, nontemporal, blocks):
    from spme_amd import engine as E
    n = (1 << 20) + 4 * 37            # not a multiple of a grid step
    print(f"case: stream probes n={n} floats blocks={blocks} nt={nontemporal}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randint(0, 3, (n,), device="cuda", generator=g).float()
    b = torch.zeros_like(a)
    E.stream_probe("copy", a, b, nontemporal=nontemporal, blocks=blocks, reps=2)
    assert torch.equal(a, b)
    E.stream_probe("write", None, b, value=2.5, nontemporal=nontemporal, blocks=blocks, reps=2)
    assert bool((b == 2.5).all())
    b.copy_(a)
    E.stream_probe("scale", None, b, value=2.0, nontemporal=nontemporal, blocks=blocks, reps=2)
    assert torch.equal(b, 8.0 * a)
    b.zero_()
    E.stream_probe("read", a, b, nontemporal=nontemporal, blocks=blocks, reps=1)
    nb = blocks if blocks else 16 * torch.cuda.get_device_properties(0).multi_processor_count
    assert float(b[:nb].double().sum()) == float(a.double().sum())
    assert bool((b[nb:] == 0).all())


def test_stream_probe_refuses_bad_arguments(torch):
    from spme_amd import engine as E
    a = torch.zeros(1000, device="cuda")
    b = torch.zeros(1000, device="cuda")
    with pytest.raises(SpmeError, match="multiple of 16"):
        E.stream_probe("copy", a[:999], b[:999])
    with pytest.raises(SpmeError, match="one float per block"):
        E.stream_probe("read", a, b, blocks=2000)
    with pytest.raises(SpmeError, match="differ in size"):
        E.stream_probe("copy", a, b[:996])
