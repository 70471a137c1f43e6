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


def test_reference_records_and_check(tmp_path):
    """the sweep's per-job check reads the reference's "#R" records and compares counts exactly and S_h to 1e-12"""
    p = tmp_path / 'nc_h0.9_N109_eps1e-10.txt'
    p.write_text('# header\n#R g2_00 S0 0.5 S1 0.5 S2 0.5 Ah 1 ref 1 prune 0 trunc 0 unres 0 nev 100 nq 3 kept 40 '
                 'skipped 0 unresdrop 0 abmiss 0 t 1.0\n#X g2_00 eps extrapolation: ...\n')
    refs = serve.reference_records(str(tmp_path))
    assert list(refs) == [('g2_00', 0.9, 109, 1e-10)]
    rec = refs[('g2_00', 0.9, 109, 1e-10)]
    good = {'S': [0.5*(1 + 5e-13), 0, 0], 'nev': 100, 'nq': 3, 'kept_top': 40}
    assert serve.check_job(good, rec)[0]
    assert not serve.check_job(dict(good, S=[0.5*(1 + 2e-12), 0, 0]), rec)[0]
    assert not serve.check_job(dict(good, kept_top=41), rec)[0]
    assert not serve.check_job(dict(good, nq=2), rec)[0]


def test_reference_records_gpu_build_counts(tmp_path):
    """a refgpu record of the same job (the reference driver's GPU build) replaces the CPU build's counts,
    not its values"""
    rec = ('#R g3_lbl S0 0.5 S1 0.5 S2 nan Ah 1 ref 1 prune 0 trunc 0 unres 0 nev %d nq 0 kept 40 '
           'skipped 7 unresdrop 0 abmiss 0 t 1.0\n')
    (tmp_path / 'nc_h0.8_N121_eps1e-12.txt').write_text(rec % 100)
    (tmp_path / 'refgpu_h0.8_N121_eps1e-12.txt').write_text((rec % 101).replace('S0 0.5', 'S0 0.6'))
    r = serve.reference_records(str(tmp_path))[('g3_lbl', 0.8, 121, 1e-12)]
    assert (r['nev'], r['S0']) == ('101', '0.5')
