"""jaxabm_amd: an MI355X-native engine for JaxABM's vectorised agent-step path."""
__version__ = "0.1.5"
