"""CPU checks of the C ABI: the library loads and exports every declared symbol."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "bpf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bpf_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported(bpf):
    names = _declared()
    assert len(names) > 30
    for n in names:
        assert hasattr(bpf.L, n), n


def test_python_signatures_cover_header(bpf):
    assert set(_declared()) == set(bpf._lib.SIGNATURES)


def test_version_and_error_string(bpf):
    assert b"gfx950" in bpf.L.bpf_version()
    assert isinstance(bpf.L.bpf_last_error(), bytes)


def test_options_round_trip_and_reject_bad_values(bpf):
    # process-wide options are host state: set, read back, restore; values
    # outside each option's range are rejected with the library's error
    import re as _re
    hdr = open(os.path.join(ROOT, "include", "bpf.h")).read()
    enum = dict((k, int(v)) for k, v in _re.findall(r"BPF_OPT_([A-Z_]+) = (\d+)", hdr))
    assert {name.upper(): code for name, code in bpf._OPTIONS.items()} == enum
    cases = {"tail_slab": ([0, 64], [32, -1]), "dense_inverse": ([0, 1, 2], [3, -1]),
             "dense_blocks": ([0, 8, 12], [-1, 257]), "tail_rounds": ([0, 2, 4], [-1, 65]),
             "wide_min": ([1, 129], [0]), "batch_threads": ([1, 4], [0, 65]), "relabel_min": ([0, 32768], [-1])}
    for name, (good, bad) in cases.items():
        old = bpf.get_option(name)
        try:
            for v in good:
                bpf.set_option(name, v)
                assert bpf.get_option(name) == v
            for v in bad:
                try:
                    bpf.set_option(name, v)
                except bpf.BpfError:
                    pass
                else:
                    raise AssertionError(f"{name}={v} accepted")
                assert bpf.get_option(name) == good[-1]
        finally:
            bpf.set_option(name, old)
