"""CPU oracle for turbosort's sort path -- TEST INFRASTRUCTURE ONLY.

``closed_form`` (numpy) and ``native`` (ctypes over the C++ restatement in
restate.cpp) are the checkers that tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg use.  Nothing in the product package imports this.
"""
