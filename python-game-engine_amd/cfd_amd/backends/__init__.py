"""Backend implementations registered behind the operator/solver registries."""
