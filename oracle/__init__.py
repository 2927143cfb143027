"""CPU oracle package (test infrastructure; see gomelt_oracle.py)."""
