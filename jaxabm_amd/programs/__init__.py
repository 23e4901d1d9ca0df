"""Library model programs: the BASELINE agent-step workloads as fused HIP
step kernels behind the reference's Model/Agent API."""
