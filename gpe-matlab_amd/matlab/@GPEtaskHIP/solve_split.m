function phi = solve_split(task,ddt,niter_inner,niter_outer)
% solve_split - lib/@GPEtask/solve_split.m:1-84 on the GPU (gpe_mex -> libgpehip.so).
%
%  Usage :
%    phi = task.solve_split(ddt,niter_inner,niter_outer)
%  Same arguments, result and side effects (current_*, history, user_callback) as
%  GPEtask.solve_split.  A static potential runs each outer iteration as one call
%  (gpe_split_outer); a time-dependent Vtd is evaluated here per block at the block's
%  start time and uploaded before that block (solve_split.m:53-54).

if(~isscalar(task.omega))
    error('GPEtaskHIP:omega', 'an array omega (Crank-Nicolson Lz step) is not supported');
end
h = task.hip;
start = task.current_iter;
if(task.Ntotal > 0)                                     % solve_split.m:17-21
    NN0 = task.Ntotal;
else
    NN0 = task.grid.integrate(abs(task.init_state).^2);
end
gpe_mex('omega', h, task.omega);
gpe_mex('params', h, task.g, task.gamma, task.decay_rate, NN0, task.mu_init, task.n_recalc, 0);
if(start>0)                                             % solve_split.m:34-38
    phi = task.current_state;
else
    phi = task.init_state;
end
gpe_mex('set_state', h, 0, phi);
gpe_mex('potential', h, 0, task.getVtotal(task.current_time));   % applyham(phi), line 41
gpe_mex('split_begin', h, ddt, niter_inner);            % lines 13-46
static = isempty(task.Vtd) && numel(task.V0) <= 1;
if(static)
    gpe_mex('potential', h, 0, task.Vtrap);
end
n_rec = task.n_recalc;
if(mod(niter_inner,n_rec) ~= 0)
    n_rec = 1;
end
dt_outer = ddt*niter_inner;
% the state crosses to the host only where MATLAB code reads it
want_phi = isa(task.user_callback,'function_handle') || task.show_image > 0;
for j=start+1:niter_outer
    if(static)
        [time2, muc, ncur] = gpe_mex('split_outer', h, j);
    else
        time=(j-1)*dt_outer;
        for jj=1:niter_inner/n_rec
            time2=time+(jj-1)*ddt*n_rec;
            gpe_mex('potential', h, 0, task.getVtotal(time2));
            [muc, ncur] = gpe_mex('split_block', h, time2);
        end
    end
    if(want_phi)
        phi = gpe_mex('get_state', h, 0);
    else
        phi = [];
    end
    task.ext_callback(phi,j,time2,muc,ncur);            % solve_split.m:80
end
phi = gpe_mex('get_state', h, 0);
task.current_state = phi;
end
