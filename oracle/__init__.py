"""CPU oracle for rLIC's line integral convolution.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product path
(``rlic_amd``) never imports, links or calls anything here.

* ``lic_oracle.cpp`` / ``liblic_oracle.so`` — C++ restatement of the
  reference's ``src/lib.rs`` (all four arithmetic profiles), wrapped by
  :mod:`oracle.cpu`.
* :mod:`oracle.pyref` — an independent pure-Python restatement with exact
  (rational) arithmetic emulating IEEE binary32/binary64, for tiny inputs.

Pinning (what checks each function against something other than itself) is
listed in each module's docstring.
"""
