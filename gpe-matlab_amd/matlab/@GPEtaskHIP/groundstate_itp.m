function [phi, varargout] = groundstate_itp(task,dt,eps,phi0)
% groundstate_itp - lib/@GPEtask/groundstate_itp.m:1-112 on the GPU.
%
% Same arguments, outputs (phi, MU, MU2, EE) and side effects (init_state,
% current_state, current_mu, current_n) as GPEtask.groundstate_itp; the iterations
% run in one library call (gpe_itp).
arguments
    task GPEtask
    dt double
    eps double
    phi0 = 'rand'
end
if(~isscalar(task.omega))
    error('GPEtaskHIP:omega', 'an array omega is not supported');
end
h = task.hip;
V = task.getVtotal(0);                                  % groundstate_itp.m:20
phi = task.process_init_state(phi0);                    % 27
task.current_state = phi;                               % 28
gpe_mex('omega', h, task.omega);
gpe_mex('params', h, task.g, 0, 0, task.Ntotal, task.mu_init, task.n_recalc, 0);
gpe_mex('potential', h, 0, V);
gpe_mex('set_state', h, 0, phi);
[MU, MU2, EE, ~, ~, converged] = gpe_mex('itp', h, dt, eps, task.itp_max_iter, task.itp_min_iter, ...
                                         task.itp_adjust_stepsize, max(nargout,1));
if(~converged)
    warning('Convergence not reached');                 % 78-81
end
phi = gpe_mex('get_state', h, 0);
task.current_state = phi;
if(nargout >= 2)                                        % 84-91
    if(task.Ntotal > 0)
        task.current_mu = MU(end);
        task.current_n = task.Ntotal;
    end
    varargout{1} = MU;
end
if(nargout >= 3)                                        % 92-100
    if(task.Ntotal > 0)
        MU2 = MU2/task.Ntotal;
    else
        MU2 = MU2./MU;
        task.current_mu = MU2(end);
        task.current_n = MU(end);
    end
    varargout{2} = MU2;
end
if(nargout >= 4)                                        % 101-108
    if(task.Ntotal > 0)
        EE = EE./task.Ntotal;
    else
        EE = EE./MU;
    end
    varargout{3} = EE;
end
task.init_state = phi;                                  % 111
end
