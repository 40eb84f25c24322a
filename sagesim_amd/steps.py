"""Ready-made step functions.

`sir_step()` / `seir_step()` are the compartment steps of the SIR/SEIR
workloads written as ordinary HIP step functions (argument order tick,
agent_index, beta, gamma[, sigma], agent_ids, breeds, locations, state).
They carry a `fused` description: when a model's only step function is one of
them, Model.setup runs the hand-written fused tick kernel instead of the
compiled source.  Both compute the same transition from the same draw, so
the result does not depend on which one runs (the tests compare them).

`reference_sir_step()` is the SIR step of the reference's own tests
(tests/test_simulation.py:24-43) in HIP: float states {1=S, 2=I, 3=R}, one
rand_uniform_philox(tick, agent_index, 1) draw per agent.
"""
from dataclasses import dataclass

from .breed import StepFunction


@dataclass(frozen=True)
class Fused:
    """Which built-in kernel computes a step: compartment model, the property
    holding the int8 state, and the globals holding its parameters."""
    kind: str                    # "sir" or "seir"
    state: str = "state"
    beta: str = "beta"
    gamma: str = "gamma"
    sigma: str = "sigma"


_SIR_SRC = r"""
// SIR {S=0, I=1, R=2}: S -> I with probability 1 - (1 - beta)^k, k = number of
// infectious neighbours, the power formed by repeated multiplication in double;
// I -> R with probability gamma.  One Philox-4x32 draw per agent and tick.
__device__ void sir_compartment_step(auto tick, auto agent_index, auto beta, auto gamma,
                                     auto agent_ids, auto breeds, auto locations, auto state) {
    const int s = (int)state[agent_index];
    const double u = rand_uniform_philox4x32(tick, agent_index);
    if (s == 0) {
        auto nb = locations[agent_index];
        int k = 0;
        for (int j = 0; j < len(nb); ++j)
            if ((int)state[nb[j]] == 1) ++k;
        if (k > 0) {
            double q = 1.0;
            const double b = 1.0 - (double)beta;
            for (int j = 0; j < k; ++j) q = q * b;
            const double p = 1.0 - q;
            if (u < p) state[agent_index] = 1;
        }
    } else if (s == 1) {
        if (u < (double)gamma) state[agent_index] = 2;
    }
}
"""

_SEIR_SRC = r"""
// SEIR {S=0, E=1, I=2, R=3}: S -> E with probability 1 - (1 - beta)^k over the
// k infectious (I) neighbours; E -> I with sigma; I -> R with gamma.
__device__ void seir_compartment_step(auto tick, auto agent_index, auto beta, auto gamma, auto sigma,
                                      auto agent_ids, auto breeds, auto locations, auto state) {
    const int s = (int)state[agent_index];
    const double u = rand_uniform_philox4x32(tick, agent_index);
    if (s == 0) {
        auto nb = locations[agent_index];
        int k = 0;
        for (int j = 0; j < len(nb); ++j)
            if ((int)state[nb[j]] == 2) ++k;
        if (k > 0) {
            double q = 1.0;
            const double b = 1.0 - (double)beta;
            for (int j = 0; j < k; ++j) q = q * b;
            const double p = 1.0 - q;
            if (u < p) state[agent_index] = 1;
        }
    } else if (s == 1) {
        if (u < (double)sigma) state[agent_index] = 2;
    } else if (s == 2) {
        if (u < (double)gamma) state[agent_index] = 3;
    }
}
"""

_REF_SIR_SRC = r"""
// tests/test_simulation.py:24-43 in HIP
__device__ void sir_step_func(auto tick, auto agent_index, auto p_infection, auto p_recovery,
                              auto agent_ids, auto breeds, auto locations, auto state_tensor) {
    auto neighbor_indices = locations[agent_index];
    const int agent_state = (int)state_tensor[agent_index];
    const double rand = rand_uniform_philox(tick, agent_index, 1);
    if (agent_state == 2 && rand < p_recovery) {
        state_tensor[agent_index] = 3;
    } else if (agent_state == 1) {
        for (int i = 0; i < len(neighbor_indices); ++i) {
            const int neighbor_state = (int)state_tensor[neighbor_indices[i]];
            if (neighbor_state == 2 && rand < p_infection) state_tensor[agent_index] = 2;
        }
    }
}
"""


def sir_step(fused=True):
    """Globals in order: beta, gamma.  Property: int8 `state`."""
    return StepFunction("sir_compartment_step", _SIR_SRC, Fused("sir") if fused else None)


def seir_step(fused=True):
    """Globals in order: beta, gamma, sigma.  Property: int8 `state`."""
    return StepFunction("seir_compartment_step", _SEIR_SRC, Fused("seir") if fused else None)


def reference_sir_step():
    """Globals in order: p_infection, p_recovery.  Property: float `state`."""
    return StepFunction("sir_step_func", _REF_SIR_SRC)
