"""NetworkSpace: the agents' graph (space.py:51-145).

Neighbours are kept per agent as a set (or, with ordered=True, an insertion
ordered list plus a set for duplicate checks), exactly as the reference
does.  Bulk edges (`connect_agents_bulk`) are kept as arrays and go straight
to the CSR builder, for graphs too large for per-edge Python calls.
"""
from typing import List, Union

import numpy as np


class NetworkSpace:
    def __init__(self, ordered: bool = False) -> None:
        self._ordered = ordered
        self._locations: list = []
        self._locations_set = {} if ordered else None
        self._bulk_u: list = []
        self._bulk_v: list = []
        self._bulk_directed: list = []
        self._agent_factory = None

    # -- agents --
    def add_agent(self, agent: int) -> None:
        agent = int(agent)
        while len(self._locations) <= agent:
            self._locations.append([] if self._ordered else set())
        if self._ordered:
            self._locations_set.setdefault(agent, set())

    def bulk_add_agents(self, total_agents: int) -> None:
        if self._ordered:
            self._locations = [[] for _ in range(total_agents)]
            self._locations_set = {i: set() for i in range(total_agents)}
        else:
            self._locations = [set() for _ in range(total_agents)]

    def get_location(self, agent_id: int) -> Union[List[int], set]:
        return self._locations[agent_id]

    def get_neighbors(self, agent_id: int) -> Union[List[int], set]:
        return self._locations[agent_id]

    @property
    def num_agents(self):
        return len(self._locations)

    # -- edges --
    def connect_agents(self, agent_0: int, agent_1: int, directed: bool = False) -> None:
        """As space.py:107-125: undirected by default; duplicates are ignored."""
        agent_0, agent_1 = int(agent_0), int(agent_1)
        if self._ordered:
            if agent_1 not in self._locations_set[agent_0]:
                self._locations_set[agent_0].add(agent_1)
                self._locations[agent_0].append(agent_1)
            if not directed and agent_0 not in self._locations_set[agent_1]:
                self._locations_set[agent_1].add(agent_0)
                self._locations[agent_1].append(agent_0)
        else:
            self._locations[agent_0].add(agent_1)
            if not directed:
                self._locations[agent_1].add(agent_0)

    def disconnect_agents(self, agent_0: int, agent_1: int, directed: bool = False) -> None:
        """As space.py:127-145."""
        agent_0, agent_1 = int(agent_0), int(agent_1)
        if self._ordered:
            if agent_1 in self._locations_set[agent_0]:
                self._locations_set[agent_0].discard(agent_1)
                self._locations[agent_0].remove(agent_1)
            if not directed and agent_0 in self._locations_set[agent_1]:
                self._locations_set[agent_1].discard(agent_0)
                self._locations[agent_1].remove(agent_0)
        else:
            self._locations[agent_0].discard(agent_1)
            if not directed:
                self._locations[agent_1].discard(agent_0)

    def connect_agents_bulk(self, u, v, directed: bool = False) -> None:
        """Many edges at once (arrays of agent ids); merged with per-edge
        connections when the CSR is built, duplicates removed."""
        u = np.asarray(u, np.int64)
        v = np.asarray(v, np.int64)
        if u.shape != v.shape:
            raise ValueError("u and v must have the same length")
        self._bulk_u.append(u)
        self._bulk_v.append(v)
        self._bulk_directed.append(bool(directed))

    def has_bulk_edges(self) -> bool:
        return bool(self._bulk_u)

    def neighbor_rows(self):
        """Per-agent neighbour containers (sets or ordered lists)."""
        return self._locations
