"""Environments on the MI355X pipeline: the reference's env layer over the
generalized step (brax/envs/base.py, brax/envs/ant.py,
brax/envs/wrappers/training.py), batched over a leading env axis."""

from brax_amd.envs.ant import Ant
from brax_amd.envs.base import Env, PipelineEnv, State
from brax_amd.envs.wrappers import AutoResetWrapper, EpisodeWrapper, Wrapper

__all__ = ["Ant", "AutoResetWrapper", "Env", "EpisodeWrapper", "PipelineEnv", "State",
           "Wrapper"]
