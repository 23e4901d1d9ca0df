"""Library model programs -- network SIR, wealth exchange, Schelling and
Deffuant -- as fused HIP step kernels behind the reference's Model/Agent
API."""
