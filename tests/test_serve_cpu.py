"""CPU tests of the serving client: the request stream is a deterministic function of its seed."""
from sincfeyn_amd import serve


def test_splitmix64_known_values():
    """splitmix64 (Steele, Lea, Flood 2014) from seed 0: the first outputs of the published generator"""
    r = serve.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_stream_deterministic_and_shaped():
    s1 = serve.make_stream(42, 10.0, 200, list(range(46)))
    s2 = serve.make_stream(42, 10.0, 200, list(range(46)))
    assert s1 == s2
    arr = [a for a, *_ in s1]
    assert all(b > a for a, b in zip(arr, arr[1:]))
    mean_gap = arr[-1]/len(arr)
    assert 0.07 < mean_gap < 0.13                       # Exp(10/s): mean 0.1 s
    assert {h for _, _, h, _, _, _ in s1} == {1.0, 0.9, 0.8, 0.7}
    assert all(N == 2*__import__('math').ceil(48/h - 1e-9) + 1 for _, _, h, N, _, _ in s1)
    assert {N for _, _, h, N, _, _ in s1} == {97, 109, 121, 139}
