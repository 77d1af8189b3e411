"""CPU oracle of the InstaHeat hot path -- test infrastructure only."""
