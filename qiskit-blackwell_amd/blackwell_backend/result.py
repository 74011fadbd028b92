"""Result objects for BlackwellSimulator.run.

When qiskit is importable the simulator returns qiskit.result.Result built
exactly as blackwell_backend/simulator.py:588-604 does; otherwise it returns
these small stand-ins with the same fields and get_counts()."""


class ExperimentResultData:
    def __init__(self, counts=None):
        self.counts = counts or {}

    def to_dict(self):
        return {"counts": self.counts}


class ExperimentResult:
    def __init__(self, shots, success, data, header, metadata=None):
        self.shots = shots
        self.success = success
        self.data = data
        self.header = header
        self.metadata = metadata or {}


class Result:
    def __init__(self, backend_name, backend_version, qobj_id, job_id, success, results,
                 time_taken=None):
        self.backend_name = backend_name
        self.backend_version = backend_version
        self.qobj_id = qobj_id
        self.job_id = job_id
        self.success = success
        self.results = results
        self.time_taken = time_taken

    def get_counts(self, experiment=0):
        return dict(self.results[experiment].data.counts)

    def data(self, experiment=0):
        return self.results[experiment].data.to_dict()


_QISKIT_RESULT = []


def _qiskit_result_classes():
    """(Result, ExperimentResult, ExperimentResultData) from qiskit, or None
    when qiskit is not installed (probed once per process)."""
    if not _QISKIT_RESULT:
        try:
            from qiskit.result import Result as QResult
            from qiskit.result.models import ExperimentResult as QExp, ExperimentResultData as QData
            _QISKIT_RESULT.append((QResult, QExp, QData))
        except ImportError:
            _QISKIT_RESULT.append(None)
    return _QISKIT_RESULT[0]


def make_result(counts, shots, header, time_taken, metadata=None):
    q = _qiskit_result_classes()
    if q is not None:
        QResult, QExp, QData = q
        kw = {"metadata": metadata} if metadata is not None else {}
        exp = QExp(shots=shots, success=True, data=QData(counts=counts), header=header, **kw)
        return QResult(backend_name="blackwell_simulator", backend_version="0.1.0", qobj_id="",
                       job_id="", success=True, results=[exp], time_taken=time_taken)
    exp = ExperimentResult(shots, True, ExperimentResultData(counts), header, metadata)
    return Result("blackwell_simulator", "0.1.0", "", "", True, [exp], time_taken)
