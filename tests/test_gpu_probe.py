"""The random-read ceiling (felog_probe_random_reads) on the device: a
positive rate for every row size, and more bytes per second for 128-byte
rows than for 16-byte ones (each random read costs at least one memory
transaction, so bigger rows deliver more bytes per transaction)."""
import pytest

import felog

pytestmark = pytest.mark.gpu


def test_random_read_probe_rates():
    rates = {}
    for rb in (16, 32, 64, 128):
        print(f"[case] probe 256 MiB table, 4M reads of {rb} B", flush=True)
        rates[rb] = felog.probe_random_reads(256 << 20, 4_000_000, rb)
        assert rates[rb] > 0
    print({k: f"{v / 1e9:.2f}G reads/s" for k, v in rates.items()})
    assert rates[128] * 128 > rates[16] * 16
