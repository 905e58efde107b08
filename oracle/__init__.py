"""CPU oracle of the D3Q19 wind-tunnel path: test infrastructure only (tests/, smoke(), bench cpu_baseline)."""
