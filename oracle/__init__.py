"""CPU oracle of the splat path (test infrastructure only; see oracle/oracle.py)."""
