"""CPU oracle of the hot path -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

The product package gridtools_amd never imports this package."""
