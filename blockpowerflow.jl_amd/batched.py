"""Many independent systems solved together."""


class BatchedSystems:
    pass
