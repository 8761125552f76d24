"""The reference pipeline's verification gates on the GPU backend."""
import pytest

pytestmark = pytest.mark.gpu


def test_verify_harness_all_gates_pass():
    from morphompm_amd import verify
    assert verify.run() == 0


def test_io_round_trip_from_device():
    from morphompm_amd import io
    assert io.main(device="cuda") == 0
