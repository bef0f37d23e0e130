"""CPU oracle (test infrastructure only; see ep_oracle.py)."""
