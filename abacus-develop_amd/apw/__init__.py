"""MI355X-native plane-wave Davidson path of ABACUS (HIP library + host mirror).

* :mod:`apw.basis`, :mod:`apw.synth`, :mod:`apw.configs` -- plane-wave basis, seeded
  synthetic Hamiltonians and the named workloads;
* :mod:`apw.pw` -- HamiltPW / DiagoDavid / device functors over libapw.so.
"""
