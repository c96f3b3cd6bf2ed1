"""GradModel: drop-in for plb/engine/function.py over the HIP engine.

``forward`` is a torch.autograd.Function whose forward runs one env step
(``substeps`` engine substeps) and returns the observation of the next step;
its backward injects the observation adjoints and runs the hand-written
adjoint substeps in reverse.  As in the reference, adjoints flow between
steps inside the engine (not through torch), the returned gradient for past
observations is zero only to keep torch's graph ordered, and steps must be
differentiated in reverse (LIFO) order -- the engine raises when they are not.
"""
import torch
from torch.autograd import Function


class GradModel:
    def __init__(self, env, softness=666., init_sampler=None, output_grid=(), return_dist=False):
        self.env = env
        self.sim = env.simulator
        self.dim = self.sim.dim
        self.primitives = self.sim.primitives
        self.substeps = self.sim.substeps
        self.controllers = env.primitives
        self.softness = softness
        self._forward_func = None
        self.output_grid = output_grid
        self.return_dist = bool(return_dist)
        if self.return_dist:
            # one column per primitive, two (sdf_2 with flag -1, +1) for the
            # 8-state ones (function.py:29-37,79-88); Chopsticks has no sdf_2
            for p in self.primitives:
                if p.state_dim == 8 and p.shape == "Chopsticks":
                    raise NotImplementedError("return_dist with Chopsticks: the reference calls sdf_2, which Chopsticks "
                                              "does not define")
        self.init_state = self.env.get_state()["state"]
        if init_sampler is None:
            init_sampler = lambda: self.init_state  # noqa: E731
        self.init_sampler = init_sampler
        self.device = "cuda"

    def reset(self, initial_states=None, device="cuda", clear_grad=True):
        """function.py:39-64."""
        self.device = device
        if initial_states is None:
            initial_states = self.init_sampler()
        self.env.set_state(initial_states, self.softness, False)
        if clear_grad:
            self.sim.clear_grads()
        return self.get_obs(0, self.device)

    def get_obs(self, s, device):
        """function.py:104-128: ([x|v] (N,6), tools (n_prim,8)[, grid_m (n,n,n)])."""
        f = s * self.sim.substeps
        x = self.sim.get_torch_xv(f).to(device)
        c = torch.zeros((len(self.primitives), 8), device=device)
        if len(self.primitives) > 0:
            c[:] = self.sim.engine.prim_state[f, :len(self.primitives)].to(device)
            for idx, p in enumerate(self.primitives):
                if p.state_dim != 8:
                    c[idx, 7] = 0
        if self.return_dist:
            x = torch.cat((x, self.sim.compute_min_dist(f).to(device)), 1)
        outputs = x.clone(), c.clone()
        if len(self.output_grid) > 0:
            for _ in self.output_grid:
                self.sim.grid_m.fill(0)
                self.sim.compute_grid_m_kernel(f)
                outputs = outputs + (self.sim.grid_m.to_torch(device),)
        return outputs

    def set_obs_grad(self, s, obs_grad, manipulator_grad, *args):
        """function.py:144-164."""
        f = s * self.sim.substeps
        if len(self.output_grid) > 0 and len(args) > 0 and args[0] is not None:
            self.sim.compute_grid_m_grad(f, args[0].detach().to("cuda", torch.float32).contiguous())
        if self.return_dist and obs_grad is not None:
            n_d = obs_grad.shape[-1] - self.dim * 2
            self.sim.compute_min_dist_grad(f, obs_grad[..., self.dim * 2:].reshape(-1, n_d))
            obs_grad = obs_grad[..., :self.dim * 2]
        obs_grad = obs_grad.reshape(-1, self.dim * 2) if obs_grad is not None else None
        mg = None
        if manipulator_grad is not None and len(self.primitives) > 0:
            mg = manipulator_grad.reshape(len(self.primitives), 8)
            mg = mg.clone()
            for idx, p in enumerate(self.primitives):
                if p.state_dim != 8:
                    mg[idx, 7] = 0
        self.sim.add_obs_grad(f, obs_grad, mg)

    def forward_step(self, s, a):
        """function.py:166-175."""
        a = a.reshape(-1).clamp(-1, 1)
        self.primitives.set_action(s, self.substeps, a)
        self.sim.engine.substep_fwd(s * self.substeps, self.substeps)

    def backward_step(self, s):
        """function.py:177-190: adjoint substeps of env step s, then the action adjoint."""
        self.sim.engine.substep_bwd(s * self.substeps, self.substeps)
        g = self.sim.engine.action_grad(s, self.substeps)
        grads = [g[j, :p.action_dim] for j, p in enumerate(self.primitives) if p.action_dim > 0]
        return torch.cat(grads).to(self.device)

    @property
    def forward(self):
        if self._forward_func is None:
            model = self

            class forward(Function):
                @staticmethod
                def forward(ctx, s, a, *past_obs):
                    ctx.save_for_backward(torch.tensor([s]), *[torch.zeros_like(i) for i in past_obs])
                    model.forward_step(s, a)
                    return model.get_obs(s + 1, a.device)

                @staticmethod
                def backward(ctx, *obs_grad):
                    tmp = ctx.saved_tensors
                    s = tmp[0].item()
                    model.set_obs_grad(s + 1, *obs_grad)
                    actor_grad = model.backward_step(s)
                    return (None, actor_grad.reshape(-1)) + tmp[1:]

            self._forward_func = forward
        return self._forward_func.apply

    def make_func(self, gamma, lam):
        """function.py:214-233: like ``forward`` with the adjoint decayed by gamma * lam."""
        model = self

        class forward(Function):
            @staticmethod
            def forward(ctx, s, a, *past_obs):
                ctx.save_for_backward(torch.tensor([s]), *[torch.zeros_like(i) for i in past_obs])
                model.forward_step(s, a)
                return model.get_obs(s + 1, a.device)

            @staticmethod
            def backward(ctx, *obs_grad):
                tmp = ctx.saved_tensors
                s = tmp[0].item()
                model.set_obs_grad(s + 1, *obs_grad)
                model.sim.scale_adjoint((s + 1) * model.substeps, gamma * lam)
                actor_grad = model.backward_step(s)
                return (None, actor_grad.reshape(-1)) + tmp[1:]

        return forward.apply
