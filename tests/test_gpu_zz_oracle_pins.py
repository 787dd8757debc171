"""The oracle's pins -- the reference's golden vectors, closed forms,
invariants and hand-worked cases, CPU only -- collected once more under the
gpu mark, so that the GPU job's record shows the checker pinned in the same
run as the parity tests that rely on it.  Each test is defined in (and runs
exactly as in) its CPU module."""
import pytest

from tests.test_oracle_closed_forms import *  # noqa: F401,F403
from tests.test_oracle_golden import *  # noqa: F401,F403
from tests.test_oracle_modes import *  # noqa: F401,F403
from tests.test_oracle_pins import *  # noqa: F401,F403
from tests.test_host_parity import *  # noqa: F401,F403
from tests.test_dataset_files import *  # noqa: F401,F403
from tests.test_datasets_cpu import *  # noqa: F401,F403
from tests.test_kernel_index_math import *  # noqa: F401,F403
from tests.test_plan_encoder import *  # noqa: F401,F403
from tests.test_exp2_fit import *  # noqa: F401,F403

pytestmark = pytest.mark.gpu
